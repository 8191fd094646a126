// Device-side helpers shared by the gfx950 kernels: error mapping, wave64 reductions,
// deterministic block partial sums, CSR row products with several lanes per row.
#pragma once

#include <hip/hip_runtime.h>

#include "../../../include/cfa_hip.h"

namespace cfa_dev {

constexpr int WAVE = 64;     // CDNA wavefront width
constexpr int BLOCK = 256;   // default workgroup: 4 waves

// Boundary tags (src/cfa/common.cuh:68-74)
constexpr int BC_INTERIOR = 0, BC_INLET = 1, BC_OUTLET = 2, BC_WALL = 3, BC_CYLINDER = 4;

// Outlet condition of a mesh (cfa_mesh_t fields; src/cfa/common.cuh:76-89)
struct OutletMode {
    int dirichlet;   // 0: p = 0 on outlet nodes; else inflow parabola on the outlet, dp pinned
    int pin;         // the node with dp = 0 when dirichlet
    double scale;    // outlet parabola factor
};

// Pressure-increment Dirichlet node (src/cfa/common.cuh:87-89 pressure_dirichlet)
__device__ __forceinline__ bool pressure_dirichlet(int i, int tag, const OutletMode& om) {
    return om.dirichlet ? i == om.pin : tag == BC_OUTLET;
}

inline hipStream_t as_stream(cfa_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline int launch_status() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? CFA_OK : (int)e;
}

inline int blocks_for(long long n, int block = BLOCK) {
    return (int)((n + block - 1) / block);
}

// Sum over a wavefront; lane 0 holds the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    return v;
}

// Sum over the workgroup in a fixed order (lanes by tree, then waves in index order); the result
// is valid in thread 0. lds: at least blockDim.x / 64 doubles.
__device__ __forceinline__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + WAVE - 1) / WAVE;
        for (int w = 0; w < nw; ++w) s += lds[w];
    }
    __syncthreads();
    return s;
}

// Sum of `count` partials (stride `stride`) by one workgroup in a fixed order; valid in all threads.
__device__ __forceinline__ double sum_partials(const double* part, int count, int stride, double* lds) {
    double s = 0.0;
    for (int k = threadIdx.x; k < count; k += blockDim.x) s += part[(size_t)k * stride];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) lds[0] = s;
    __syncthreads();
    s = lds[0];
    __syncthreads();
    return s;
}

// K sums of partials at once, each in the order of sum_partials (lane-strided, the wave tree, the
// waves in order), so the values are bit-identical to K calls of it, with one pair of barriers
// instead of K: out[q] = sum of parts[q][k * stride], k < count[q]; valid in all threads.
// lds: at least K * blockDim.x / 64 doubles.
template <int K>
__device__ __forceinline__ void sum_partials_k(const double* const (&parts)[K], const int (&count)[K], int stride,
                                               double (&out)[K], double* lds) {
    double s[K];
    int cmax = 0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        s[q] = 0.0;
        cmax = count[q] > cmax ? count[q] : cmax;
    }
    for (int k = threadIdx.x; k < cmax; k += blockDim.x)
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (k < count[q]) s[q] += parts[q][(size_t)k * stride];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        s[q] = wave_sum(s[q]);
        if (lane == 0) lds[q * nw + wid] = s[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += lds[q * nw + w];
        out[q] = t;
    }
    __syncthreads();
}

// K block sums at once (block_sum's order: the wave tree, then the waves in order); valid in
// thread 0. lds: at least K * blockDim.x / 64 doubles.
template <int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], double* lds) {
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        v[q] = wave_sum(v[q]);
        if (lane == 0) lds[q * nw + wid] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) {
            double t = 0.0;
            for (int w = 0; w < nw; ++w) t += lds[q * nw + w];
            v[q] = t;
        }
    __syncthreads();
}

// Sum over TPR consecutive lanes (TPR a power of two <= 64); lane 0 of the group holds it.
template <int TPR>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_down(v, o, TPR);
    return v;
}

template <int TPR>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = TPR / 2; o > 0; o >>= 1) v += __shfl_down(v, o, TPR);
    return v;
}

// The lane's entries of the CSR row [k0, k1) (k0 + lane, k0 + lane + TPR, ...), ROW_UNROLL at a
// time: the column indices and values of a group are all loaded before any is used, then f(a, j)
// runs for each entry in order. Slots past the row's end give (0, column 0), so every gather f
// makes is unconditional and the group's gathers are in flight together; an accumulation of
// a * x[j] adds an exact zero for them. (Non-temporal loads of the matrix were slower: the lanes
// of a row share cache lines across the group's load instructions.)
constexpr int ROW_UNROLL = 4;
template <int TPR, typename T, typename F>
__device__ __forceinline__ void row_entries(const int* __restrict__ ci, const T* __restrict__ val, int k0, int k1,
                                            int lane, F f) {
    for (int k = k0 + lane; k < k1; k += ROW_UNROLL * TPR) {
        int c[ROW_UNROLL];
        T v[ROW_UNROLL];
#pragma unroll
        for (int u = 0; u < ROW_UNROLL; ++u) {
            const int idx = k + u * TPR;
            const bool ok = idx < k1;
            c[u] = ok ? ci[idx] : 0;
            v[u] = ok ? val[idx] : T(0);
        }
#pragma unroll
        for (int u = 0; u < ROW_UNROLL; ++u) f(v[u], c[u]);
    }
}

// row_entries over NV value arrays on one pattern (the OSS operator's K, Bx, By): f(v, j) gets the
// entry's NV values v[0..NV-1] and its column; padded slots give zeros and column 0.
template <int NV, typename T>
struct RowValues {
    const T* p[NV];
};
template <int TPR, int NV, typename T, typename F>
__device__ __forceinline__ void row_entries_n(const int* __restrict__ ci, const RowValues<NV, T>& vals, int k0,
                                              int k1, int lane, F f) {
    for (int k = k0 + lane; k < k1; k += ROW_UNROLL * TPR) {
        int c[ROW_UNROLL];
        T v[ROW_UNROLL][NV];
#pragma unroll
        for (int u = 0; u < ROW_UNROLL; ++u) {
            const int idx = k + u * TPR;
            const bool ok = idx < k1;
            c[u] = ok ? ci[idx] : 0;
#pragma unroll
            for (int m = 0; m < NV; ++m) v[u][m] = ok ? vals.p[m][idx] : T(0);
        }
#pragma unroll
        for (int u = 0; u < ROW_UNROLL; ++u) f(v[u], c[u]);
    }
}

// Row product sum_k val[k] x[ci[k]] over [s, e) with TPR lanes (lane = rank inside the group)
template <int TPR>
__device__ __forceinline__ double row_dot(const int* __restrict__ ci, const double* __restrict__ val,
                                          const double* __restrict__ x, int s, int e, int lane) {
    double acc = 0.0;
    row_entries<TPR>(ci, val, s, e, lane, [&](double a, int j) { acc += a * x[j]; });
    return group_sum<TPR>(acc);
}

// Looks up, once per device, the device address of the fixed workspace of the no-allocation
// entry point (fem.hip), for the current device; cfa_init calls it.
int scratch_setup();

}  // namespace cfa_dev
