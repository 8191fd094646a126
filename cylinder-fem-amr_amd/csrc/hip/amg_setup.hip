// SA-AMG set-up on gfx950: the numeric phases of src/cfa/linalg.cuh:995-1060 and 1276-1431
// (AmgPreconditioner::setup, spectral_radius_DinvA, smoothed_prolongator, csr_multiply,
// csr_transpose), which the reference runs on the host with OpenMP (the greedy aggregation is in
// aggregate.hip). Every phase reproduces the host's floating-point results bit for bit:
//   * the power iteration keeps the host's chunking (rows in chunks of 4096, each chunk summed in
//     row order, the chunk sums added in chunk order);
//   * a sparse product accumulates every output entry in the host's encounter order (row entries
//     of the left factor in order, each right row in order), starting from 0.0;
//   * contraction is off, so every product and sum is the host's own operation.
// Nothing is summed by atomics: products and P are built row by row (rows beyond the LDS kernel's
// 4096 products through a stable sort of the listed products in global memory), the transpose is
// a stable radix sort (rocPRIM) of P's entries by column that moves values without combining them.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "cfa_device.hpp"

using namespace cfa_dev;

#pragma clang fp contract(off)

namespace {

constexpr int CHUNK = 4096;          // host chunk of spectral_radius_DinvA
constexpr int SPGEMM_CAP = 4096;     // products per output row held in LDS
constexpr int SPGEMM_THREADS = 256;

// inverse diagonal (1 where |a_ii| <= 1e-300, the AMG rule of linalg.cuh:871-878)
__global__ void k_amg_dinv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int a = rp[i]; a < rp[i + 1]; ++a)
        if (ci[a] == i) d = v[a];
    dinv[i] = fabs(d) > 1.0e-300 ? 1.0 / d : 1.0;
}

// Sum of cnt values in index order by thread 0 (t += v[c], c = 0, 1, ...), the values staged in
// LDS by the whole workgroup a tile at a time, so the loads are in flight together instead of one
// dependent global load per term. All threads call it; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double ordered_sum(const double* __restrict__ v, int cnt, double* tile) {
    double t = 0.0;
    for (int c0 = 0; c0 < cnt; c0 += NT) {
        const int m = min(NT, cnt - c0);
        if ((int)threadIdx.x < m) tile[threadIdx.x] = v[c0 + threadIdx.x];
        __syncthreads();
        if (threadIdx.x == 0)
#pragma unroll 8
            for (int c = 0; c < m; ++c) t += tile[c];
        __syncthreads();
    }
    return t;
}

// One power step y = D^-1 A x, one thread per row, entries in row order (loaded ahead of use).
// The step's x is formed where it is read: x0 on the first step, else y_prev / nrm with
// nrm = sqrt(sum of the previous step's chunk sums, in chunk order), which every workgroup
// recomputes (the same division the host applies to its stored x). Each row also stores its own
// x_i for the chunk sums.
__global__ void __launch_bounds__(BLOCK)
k_power_step(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
             const double* __restrict__ dinv, const double* __restrict__ x0, const double* __restrict__ y_prev,
             const double* __restrict__ py_prev, int nch, int first, double* __restrict__ x,
             double* __restrict__ y) {
    __shared__ double tile[BLOCK];
    __shared__ double s_nrm;
    if (!first) {
        const double t = ordered_sum<BLOCK>(py_prev, nch, tile);
        if (threadIdx.x == 0) s_nrm = sqrt(t);
        __syncthreads();
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double nrm = first ? 1.0 : s_nrm;
    double s = 0.0;
    row_entries<1>(ci, v, rp[i], rp[i + 1], 0, [&](double a, int j) {
        const double xj = first ? x0[j] : y_prev[j] / nrm;
        s += a * xj;
    });
    x[i] = first ? x0[i] : y_prev[i] / nrm;
    y[i] = dinv[i] * s;
}

// per chunk of 4096 rows: sum y_i^2 and x_i^2 in row order. A workgroup of 1024 threads per chunk
// stages the squares in LDS; then lane 0 adds the y^2 and lane 1 the x^2 in row order, two
// dependent chains running side by side in one instruction stream with their LDS reads issued
// ahead.
constexpr int CHUNK_THREADS = 1024;
__global__ void __launch_bounds__(CHUNK_THREADS) k_power_chunks(int n, const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                double* __restrict__ py, double* __restrict__ px) {
    __shared__ double sq[2][CHUNK];
    const int c = blockIdx.x, t = threadIdx.x;
    const int lo = c * CHUNK, cnt = min(n, lo + CHUNK) - lo;
#pragma unroll 4
    for (int k = t; k < cnt; k += CHUNK_THREADS) {
        const double yi = y[lo + k], xi = x[lo + k];
        sq[0][k] = yi * yi;
        sq[1][k] = xi * xi;
    }
    __syncthreads();
    if (t < 2) {
        const double* q = sq[t];
        double acc = 0.0;
#pragma unroll 16
        for (int k = 0; k < cnt; ++k) acc += q[k];
        if (t == 0) py[c] = acc;
        else px[c] = acc;
    }
}

// nrm = sqrt(sum of chunk sums), lambda = nrm / sqrt(sum of x^2), in chunk order; out[0] = nrm,
// out[1] = lambda
__global__ void __launch_bounds__(BLOCK) k_power_final(int nch, const double* __restrict__ py,
                                                       const double* __restrict__ px, double* __restrict__ out) {
    __shared__ double tile[BLOCK];
    double nrm = ordered_sum<BLOCK>(py, nch, tile);
    const double xn = ordered_sum<BLOCK>(px, nch, tile);
    if (threadIdx.x != 0) return;
    nrm = sqrt(nrm);
    out[0] = nrm;
    out[1] = nrm / sqrt(xn);
}

// ---- sparse product C = A B, host csr_multiply semantics ------------------------------------------
// Every row of B has sorted, unique columns, so the host's encounter order for an output column
// j is the order of the entries of A's row: c_ij = sum over a of a_ia b_(k_a) j, in a's order,
// from 0. One thread per row merges the B rows of its A entries in ascending column order, with
// the cursors, the current head columns and the A values in registers (the loops over the
// MERGE_MAX slots are unrolled): each output column costs one round of independent loads (the
// values of the hits and the next heads). Rows of A longer than MERGE_MAX entries set
// *overflow = 1 (the caller then uses k_spgemm_rows).
constexpr int MERGE_MAX = 16;
constexpr int NO_COL = 0x7fffffff;

__global__ void __launch_bounds__(BLOCK)
k_spgemm_merge(int rows, const int* __restrict__ a_rp, const int* __restrict__ a_ci, const double* __restrict__ a_v,
               const int* __restrict__ b_rp, const int* __restrict__ b_ci, const double* __restrict__ b_v, int pass,
               int* __restrict__ cnt, const int* __restrict__ c_rp, int* __restrict__ c_ci, double* __restrict__ c_v,
               int* __restrict__ overflow) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const int a0 = a_rp[row], len = a_rp[row + 1] - a0;
    if (len > MERGE_MAX) {
        overflow[0] = 1;
        if (pass == 0) cnt[row] = 0;
        return;
    }
    int cur[MERGE_MAX], end[MERGE_MAX], head[MERGE_MAX];
    double av[MERGE_MAX];
#pragma unroll
    for (int t = 0; t < MERGE_MAX; ++t) {
        cur[t] = 0;
        end[t] = 0;
        av[t] = 0.0;
        if (t < len) {
            const int k = a_ci[a0 + t];
            av[t] = a_v[a0 + t];
            cur[t] = b_rp[k];
            end[t] = b_rp[k + 1];
        }
    }
#pragma unroll
    for (int t = 0; t < MERGE_MAX; ++t) head[t] = cur[t] < end[t] ? b_ci[cur[t]] : NO_COL;
    int out = 0;
    const int o0 = pass == 1 ? c_rp[row] : 0;
    while (true) {
        int j = NO_COL;
#pragma unroll
        for (int t = 0; t < MERGE_MAX; ++t) j = min(j, head[t]);
        if (j == NO_COL) break;
        if (pass == 1) {
            double bv[MERGE_MAX];
#pragma unroll
            for (int t = 0; t < MERGE_MAX; ++t) bv[t] = head[t] == j ? b_v[cur[t]] : 0.0;
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < MERGE_MAX; ++t)
                if (head[t] == j) {
                    const double prod = av[t] * bv[t];
                    acc += prod;
                }
            c_ci[o0 + out] = j;
            c_v[o0 + out] = acc;
        }
#pragma unroll
        for (int t = 0; t < MERGE_MAX; ++t)
            if (head[t] == j) {
                ++cur[t];
                head[t] = cur[t] < end[t] ? b_ci[cur[t]] : NO_COL;
            }
        ++out;
    }
    if (pass == 0) cnt[row] = out;
}

// Rows of A with up to 64 entries and up to WMERGE_CAP products: one wavefront per row. Lane l
// holds entry l of A's row; the wave stages the B rows in LDS (columns, and in pass 1 the
// products a_il b_lj rounded as the host rounds them) with independent loads, then merges them:
// per output column j the wave takes the minimum head column, and the lanes whose head is j add
// their products from 0.0 in lane order, which is A's entry order (the host's encounter order).
// Each output column costs a few wave operations on LDS instead of the bitonic sort's stages.
// Output entries are buffered one per lane and written 64 at a time.
constexpr int WMERGE_CAP = 1024;

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
    return v;
}

// v of lane l (l wave-uniform)
__device__ __forceinline__ double read_lane(double v, int l) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__global__ void __launch_bounds__(WAVE)
k_spgemm_wave(int rows, const int* __restrict__ a_rp, const int* __restrict__ a_ci, const double* __restrict__ a_v,
              const int* __restrict__ b_rp, const int* __restrict__ b_ci, const double* __restrict__ b_v, int pass,
              int* __restrict__ cnt, const int* __restrict__ c_rp, int* __restrict__ c_ci, double* __restrict__ c_v,
              int* __restrict__ overflow) {
    __shared__ int s_col[WMERGE_CAP];
    __shared__ double s_prod[WMERGE_CAP];
    __shared__ int s_off[WAVE + 1], s_b0[WAVE];
    __shared__ double s_av[WAVE];
    const int row = blockIdx.x, lane = threadIdx.x;
    if (row >= rows) return;
    const int a0 = a_rp[row], len = a_rp[row + 1] - a0;
    if (len > WAVE) {
        if (lane == 0) {
            overflow[0] = 1;
            if (pass == 0) cnt[row] = 0;
        }
        return;
    }
    int b0 = 0, blen = 0;
    double av = 0.0;
    if (lane < len) {
        const int k = a_ci[a0 + lane];
        av = a_v[a0 + lane];
        b0 = b_rp[k];
        blen = b_rp[k + 1] - b0;
    }
    // inclusive scan of the B row lengths over the wave
    int incl = blen;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int t = __shfl_up(incl, o, WAVE);
        if (lane >= o) incl += t;
    }
    const int total = __shfl(incl, WAVE - 1, WAVE);
    if (total > WMERGE_CAP) {
        if (lane == 0) {
            overflow[0] = 1;
            if (pass == 0) cnt[row] = 0;
        }
        return;
    }
    const int off = incl - blen;
    s_off[lane] = off;
    if (lane == WAVE - 1) s_off[WAVE] = total;
    s_b0[lane] = b0;
    s_av[lane] = av;
    __syncthreads();
    // stage: product p belongs to the lane whose range [s_off[l], s_off[l + 1]) holds it
    for (int p = lane; p < total; p += WAVE) {
        int lo = 0, hi = WAVE - 1;   // the last l with s_off[l] <= p
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        const int src = s_b0[lo] + (p - s_off[lo]);
        s_col[p] = b_ci[src];
        if (pass == 1) {
            const double prod = s_av[lo] * b_v[src];
            s_prod[p] = prod;
        }
    }
    __syncthreads();
    int cur = off;
    const int end = off + blen;
    int head = cur < end ? s_col[cur] : NO_COL;
    int out = 0, my_j = 0;
    double my_v = 0.0;
    const int o0 = pass == 1 ? c_rp[row] : 0;
    while (true) {
        const int j = wave_min_int(head);
        if (j == NO_COL) break;
        const bool hit = head == j;
        if (pass == 1) {
            const double prod = hit ? s_prod[cur] : 0.0;
            unsigned long long mask = __ballot(hit);
            double acc = 0.0;
            while (mask) {
                const int l = __ffsll((long long)mask) - 1;
                acc += read_lane(prod, l);
                mask &= mask - 1;
            }
            const int slot = out & (WAVE - 1);
            if (lane == slot) {
                my_j = j;
                my_v = acc;
            }
            if (slot == WAVE - 1) {
                c_ci[o0 + out - (WAVE - 1) + lane] = my_j;
                c_v[o0 + out - (WAVE - 1) + lane] = my_v;
            }
        }
        if (hit) {
            ++cur;
            head = cur < end ? s_col[cur] : NO_COL;
        }
        ++out;
    }
    if (pass == 1) {
        const int rem = out & (WAVE - 1);
        if (lane < rem) {
            c_ci[o0 + out - rem + lane] = my_j;
            c_v[o0 + out - rem + lane] = my_v;
        }
    }
    if (pass == 0 && lane == 0) cnt[row] = out;
}

// Products of long rows: one workgroup per row of A. The products a_ik b_kj are listed in encounter order (entries of
// A's row in order, each B row in order) into LDS with key (j, position); a bitonic sort by key
// groups equal columns with positions ascending; each column is then summed from 0.0 in that
// order. Pass 0 writes the number of distinct columns of each row to cnt[row] (and sets *overflow
// when a row has more than SPGEMM_CAP products); pass 1 writes columns and values at c_rp[row].
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_tmp, int& total) {
    // simple two-level scan over the workgroup (SPGEMM_THREADS lanes)
    const int t = threadIdx.x;
    s_tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < SPGEMM_THREADS; off <<= 1) {
        const int add = t >= off ? s_tmp[t - off] : 0;
        __syncthreads();
        s_tmp[t] += add;
        __syncthreads();
    }
    const int incl = s_tmp[t];
    total = s_tmp[SPGEMM_THREADS - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(SPGEMM_THREADS)
k_spgemm_rows(int rows, const int* __restrict__ a_rp, const int* __restrict__ a_ci, const double* __restrict__ a_v,
              const int* __restrict__ b_rp, const int* __restrict__ b_ci, const double* __restrict__ b_v, int pass,
              int* __restrict__ cnt, const int* __restrict__ c_rp, int* __restrict__ c_ci, double* __restrict__ c_v,
              int* __restrict__ overflow) {
    __shared__ unsigned long long key[SPGEMM_CAP];
    __shared__ double val[SPGEMM_CAP];
    __shared__ int s_tmp[SPGEMM_THREADS];
    const int row = blockIdx.x;
    if (row >= rows) return;
    const int t = threadIdx.x;
    const int a0 = a_rp[row], a1 = a_rp[row + 1];
    // list the products in encounter order
    int filled = 0;
    bool over = false;
    for (int base = a0; base < a1; base += SPGEMM_THREADS) {
        const int a = base + t;
        const int k = a < a1 ? a_ci[a] : 0;
        const int len = a < a1 ? b_rp[k + 1] - b_rp[k] : 0;
        int total = 0;
        const int off = block_exclusive_scan(len, s_tmp, total);
        if (filled + total > SPGEMM_CAP) {
            over = true;
            break;
        }
        if (a < a1) {
            const double av = a_v[a];
            const int b0 = b_rp[k];
            for (int q = 0; q < len; ++q) {
                const int pos = filled + off + q;
                key[pos] = ((unsigned long long)(unsigned)b_ci[b0 + q] << 32) | (unsigned)pos;
                val[pos] = av * b_v[b0 + q];
            }
        }
        filled += total;
        __syncthreads();
    }
    if (over) {
        if (t == 0) {
            overflow[0] = 1;
            if (pass == 0) cnt[row] = 0;
        }
        return;
    }
    // bitonic sort of the keys (padded to a power of two with the largest key)
    int size = 1;
    while (size < filled) size <<= 1;
    for (int p = filled + t; p < size; p += SPGEMM_THREADS) key[p] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p = t; p < size; p += SPGEMM_THREADS) {
                const int q = p ^ j;
                if (q > p) {
                    const unsigned long long kp = key[p], kq = key[q];
                    const bool up = (p & k) == 0;
                    if ((kp > kq) == up) {
                        key[p] = kq;
                        key[q] = kp;
                    }
                }
            }
            __syncthreads();
        }
    // segment heads: first entry of each distinct column
    int distinct_before = 0;
    int n_distinct = 0;
    for (int base = 0; base < filled; base += SPGEMM_THREADS) {
        const int p = base + t;
        const int head = p < filled && (p == 0 || (key[p] >> 32) != (key[p - 1] >> 32)) ? 1 : 0;
        int total = 0;
        const int off = block_exclusive_scan(head, s_tmp, total);
        if (pass == 1 && head) {
            // sum this column's products in encounter order (positions ascending)
            const unsigned col = (unsigned)(key[p] >> 32);
            double acc = 0.0;
            for (int r = p; r < filled && (unsigned)(key[r] >> 32) == col; ++r)
                acc += val[(unsigned)(key[r] & 0xffffffffull)];
            const int o = c_rp[row] + distinct_before + off;
            c_ci[o] = (int)col;
            c_v[o] = acc;
        }
        distinct_before += total;
        n_distinct += total;
    }
    if (pass == 0 && t == 0) cnt[row] = n_distinct;
}

// exclusive scan of cnt[0..n) into out[0..n] (out[n] = total): rocPRIM's device scan over n + 1
// values (cnt, then 0); integer sums, so the result does not depend on its order
struct CountAt {
    const int* c;
    int n;
    __host__ __device__ int operator()(int i) const { return i < n ? c[i] : 0; }
};

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) {
    if (blockIdx.x == 0 && threadIdx.x == 0) dst[0] = src[0];
}

size_t scan_bytes(int n) {
    size_t bytes = 0;
    const auto in = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0), CountAt{nullptr, n});
    if (rocprim::exclusive_scan(nullptr, bytes, in, static_cast<int*>(nullptr), 0, (size_t)n + 1,
                                rocprim::plus<int>()) != hipSuccess)
        return 0;
    return bytes;
}

// ---- smoothed prolongator P = P0 - omega D^-1 A P0 (linalg.cuh smoothed_prolongator): per row,
// AP0's entries in column order, the row's own aggregate merged in, exact zeros dropped
__global__ void k_prolongator_rows(int n, const int* __restrict__ ap_rp, const int* __restrict__ ap_ci,
                                   const double* __restrict__ ap_v, const int* __restrict__ agg,
                                   const double* __restrict__ p0v, const double* __restrict__ dinv, double omega,
                                   int pass, int* __restrict__ cnt, const int* __restrict__ p_rp,
                                   int* __restrict__ p_ci, double* __restrict__ p_v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int out = 0;
    bool own = false;
    const int oc = agg[i];
    const double own_v = p0v[i];
    const int o0 = pass == 1 ? p_rp[i] : 0;
    for (int a = ap_rp[i]; a < ap_rp[i + 1]; ++a) {
        const int j = ap_ci[a];
        if (!own && j > oc) {
            if (pass == 1) { p_ci[o0 + out] = oc; p_v[o0 + out] = own_v; }
            ++out;
            own = true;
        }
        double val = -omega * dinv[i] * ap_v[a];
        if (j == oc) { val += own_v; own = true; }
        if (val != 0.0) {
            if (pass == 1) { p_ci[o0 + out] = j; p_v[o0 + out] = val; }
            ++out;
        }
    }
    if (!own) {
        if (pass == 1) { p_ci[o0 + out] = oc; p_v[o0 + out] = own_v; }
        ++out;
    }
    if (pass == 0) cnt[i] = out;
}

// ---- R = P^T (csr_transpose: entries of each column in ascending row order) ----------------------
// ---- transpose (host csr_transpose: the entries of a column in ascending row order) ----------------
// A stable radix sort of the entries by column, from their CSR order (rows ascending), lists
// every column's entries in row order; the values are moved, never combined.
__global__ void k_row_of_entries(int rows, const int* __restrict__ rp, int* __restrict__ row_of) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    for (int k = rp[i]; k < rp[i + 1]; ++k) row_of[k] = i;
}

__global__ void k_transpose_gather(int nnz, const int* __restrict__ order, const int* __restrict__ row_of,
                                   const double* __restrict__ v, int* __restrict__ t_ci, double* __restrict__ t_v) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int k = order[p];
    t_ci[p] = row_of[k];
    t_v[p] = v[k];
}

// t_rp[c] = first sorted position whose column is >= c, for c = 0..cols
__global__ void k_column_starts(int cols, int nnz, const unsigned* __restrict__ keys, int* __restrict__ t_rp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > cols) return;
    int lo = 0, hi = nnz;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (unsigned)c) lo = mid + 1;
        else hi = mid;
    }
    t_rp[c] = lo;
}

unsigned column_bits(int cols) {
    unsigned b = 1;
    while (b < 31 && (1u << b) < (unsigned)cols) ++b;
    return b;
}

size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

size_t transpose_sort_bytes(int nnz, int cols) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
                                  rocprim::counting_iterator<int>(0), static_cast<int*>(nullptr), (size_t)nnz, 0,
                                  column_bits(cols)) != hipSuccess)
        return 0;
    return bytes;
}

int blocks(int n) { return blocks_for(n > 0 ? n : 1); }

// ---- products beyond the LDS method: expand, stable sort, compress -----------------------------
// Every product is listed in encounter order (the left row's entries in order, each right row in
// order) under the key (row, column); a stable radix sort groups equal keys and keeps each group
// in that order, and each group is summed from 0.0 in order: the host's accumulation order.
struct CountAt64 {
    const long long* c;
    int n;
    __host__ __device__ long long operator()(int i) const { return i < n ? c[i] : 0; }
};

__global__ void k_entry_products(int nnz, const int* __restrict__ a_ci, const int* __restrict__ b_rp,
                                 long long* __restrict__ cnt) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnz) return;
    const int k = a_ci[a];
    cnt[a] = b_rp[k + 1] - b_rp[k];
}

__global__ void k_expand_products(int nnz, const int* __restrict__ row_of, const int* __restrict__ a_ci,
                                  const double* __restrict__ a_v, const int* __restrict__ b_rp,
                                  const int* __restrict__ b_ci, const double* __restrict__ b_v,
                                  const long long* __restrict__ off, unsigned cb, unsigned long long* __restrict__ key,
                                  double* __restrict__ val) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nnz) return;
    const int k = a_ci[a];
    const unsigned long long r = (unsigned long long)(unsigned)row_of[a] << cb;
    const double av = a_v[a];
    long long o = off[a];
    for (int q = b_rp[k]; q < b_rp[k + 1]; ++q, ++o) {
        key[o] = r | (unsigned)b_ci[q];
        val[o] = av * b_v[q];
    }
}

__global__ void k_group_heads(long long n, const unsigned long long* __restrict__ key, int* __restrict__ head) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    head[p] = (p == 0 || key[p] != key[p - 1]) ? 1 : 0;
}

__global__ void k_group_sums(long long n, const unsigned long long* __restrict__ key, const double* __restrict__ val,
                             const int* __restrict__ gidx, unsigned cb, int* __restrict__ c_ci,
                             double* __restrict__ c_v, int* __restrict__ grow) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || (p > 0 && key[p] == key[p - 1])) return;
    const unsigned long long k = key[p];
    double acc = 0.0;
    for (long long q = p; q < n && key[q] == k; ++q) acc += val[q];
    const int g = gidx[p];
    c_ci[g] = (int)(k & ((1ull << cb) - 1ull));
    c_v[g] = acc;
    grow[g] = (int)(k >> cb);
}

// c_rp[i] = first group whose row is >= i, for i = 0..rows (grow ascending)
__global__ void k_row_starts(int rows, int groups, const int* __restrict__ grow, int* __restrict__ c_rp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > rows) return;
    int lo = 0, hi = groups;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (grow[mid] < i) lo = mid + 1;
        else hi = mid;
    }
    c_rp[i] = lo;
}

// workspace of the sorted product, in 256-byte aligned pieces
struct SortedWs {
    long long* cnt;            // [a_nnz] products of each left entry
    long long* off;            // [a_nnz + 1] their offsets (off[a_nnz] = products)
    int* row_of;               // [a_nnz]
    unsigned long long* key[2];   // [products] listed, sorted
    double* val[2];            // [products] listed, sorted
    int* head;                 // [products] group heads, then the group rows
    int* gidx;                 // [products + 1] group of each head
    void* tmp;
    size_t tmp_bytes;
};

size_t sorted_tmp_bytes(int a_nnz, long long products, unsigned end_bit) {
    size_t need = 0, b = 0;
    const auto in64 = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0), CountAt64{nullptr, a_nnz});
    if (rocprim::exclusive_scan(nullptr, b, in64, static_cast<long long*>(nullptr), 0LL, (size_t)a_nnz + 1,
                                rocprim::plus<long long>()) != hipSuccess)
        return 0;
    need = b;
    if (products > 0) {
        b = 0;
        if (rocprim::radix_sort_pairs(nullptr, b, static_cast<const unsigned long long*>(nullptr),
                                      static_cast<unsigned long long*>(nullptr), static_cast<const double*>(nullptr),
                                      static_cast<double*>(nullptr), (size_t)products, 0, end_bit) != hipSuccess)
            return 0;
        need = b > need ? b : need;
        b = 0;
        const auto inh = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0),
                                                          CountAt{nullptr, (int)products});
        if (rocprim::exclusive_scan(nullptr, b, inh, static_cast<int*>(nullptr), 0, (size_t)products + 1,
                                    rocprim::plus<int>()) != hipSuccess)
            return 0;
        need = b > need ? b : need;
    }
    return need;
}

size_t sorted_layout(int a_nnz, long long products, unsigned end_bit, void* base, SortedWs* w) {
    char* p = static_cast<char*>(base);
    size_t at = 0;
    auto take = [&](size_t bytes) {
        char* q = p ? p + at : nullptr;
        at += align_up(bytes > 0 ? bytes : 1);
        return q;
    };
    const size_t P = (size_t)products;
    SortedWs s;
    s.cnt = reinterpret_cast<long long*>(take((size_t)a_nnz * sizeof(long long)));
    s.off = reinterpret_cast<long long*>(take(((size_t)a_nnz + 1) * sizeof(long long)));
    s.row_of = reinterpret_cast<int*>(take((size_t)a_nnz * sizeof(int)));
    s.key[0] = reinterpret_cast<unsigned long long*>(take(P * sizeof(unsigned long long)));
    s.key[1] = reinterpret_cast<unsigned long long*>(take(P * sizeof(unsigned long long)));
    s.val[0] = reinterpret_cast<double*>(take(P * sizeof(double)));
    s.val[1] = reinterpret_cast<double*>(take(P * sizeof(double)));
    s.head = reinterpret_cast<int*>(take(P * sizeof(int)));
    s.gidx = reinterpret_cast<int*>(take((P + 1) * sizeof(int)));
    s.tmp_bytes = sorted_tmp_bytes(a_nnz, products, end_bit);
    s.tmp = take(s.tmp_bytes);
    if (w) *w = s;
    return s.tmp_bytes ? at : 0;
}

}  // namespace

extern "C" {

int cfa_amg_dinv(int n, const int* rp, const int* ci, const double* val, double* dinv, cfa_stream_t stream) {
    if (n < 0 || (n > 0 && (!rp || !ci || !val || !dinv))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_amg_dinv, dim3(blocks(n)), dim3(BLOCK), 0, as_stream(stream), n, rp, ci, val, dinv);
    return launch_status();
}

namespace {
// power-iteration workspace: x [n], y [2][n], py [2][nch], px [nch], result {nrm, lambda}
struct PowerWs {
    double *x, *y[2], *py[2], *px, *out;
};
PowerWs power_ws(double* ws, int n) {
    const int nch = (n + CHUNK - 1) / CHUNK;
    PowerWs w;
    w.x = ws;
    w.y[0] = ws + n;
    w.y[1] = ws + 2 * (size_t)n;
    w.py[0] = ws + 3 * (size_t)n;
    w.py[1] = w.py[0] + nch;
    w.px = w.py[1] + nch;
    w.out = w.px + nch;
    return w;
}
}  // namespace

size_t cfa_amg_power_ws_doubles(int n) {
    return n < 0 ? 0 : 3 * (size_t)n + 3 * (size_t)((n + CHUNK - 1) / CHUNK) + 2;
}

int cfa_amg_power(int n, const int* rp, const int* ci, const double* val, const double* dinv, const double* x0,
                  int iterations, double* ws, double* lambda, cfa_stream_t stream) {
    if (n <= 0 || !rp || !ci || !val || !dinv || !x0 || !ws || iterations < 1) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    const int nch = (n + CHUNK - 1) / CHUNK;
    const PowerWs w = power_ws(ws, n);
    for (int it = 0; it < iterations; ++it) {
        const int cur = it & 1, prev = cur ^ 1;
        hipLaunchKernelGGL(k_power_step, dim3(blocks(n)), dim3(BLOCK), 0, s, n, rp, ci, val, dinv, x0, w.y[prev],
                           w.py[prev], nch, it == 0 ? 1 : 0, w.x, w.y[cur]);
        hipLaunchKernelGGL(k_power_chunks, dim3(nch), dim3(CHUNK_THREADS), 0, s, n, w.x, w.y[cur], w.py[cur], w.px);
    }
    hipLaunchKernelGGL(k_power_final, dim3(1), dim3(BLOCK), 0, s, nch, w.py[(iterations - 1) & 1], w.px, w.out);
    int rc = launch_status();
    if (rc || !lambda) return rc;
    hipError_t e = hipMemcpyAsync(lambda, w.out + 1, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CFA_OK : (int)e;
}

int cfa_spgemm(int rows, const int* a_rp, const int* a_ci, const double* a_v, const int* b_rp, const int* b_ci,
               const double* b_v, int pass, int method, int* cnt, const int* c_rp, int* c_ci, double* c_v,
               int* overflow, cfa_stream_t stream) {
    if (rows < 0 || (pass != 0 && pass != 1) || method < 0 || method > 2 || !overflow) return CFA_E_ARG;
    if (rows == 0) return CFA_OK;
    if (!a_rp || !a_ci || !a_v || !b_rp || !b_ci || !b_v) return CFA_E_ARG;
    if ((pass == 0 && !cnt) || (pass == 1 && (!c_rp || !c_ci || !c_v))) return CFA_E_ARG;
    if (method == 0)
        hipLaunchKernelGGL(k_spgemm_merge, dim3(blocks(rows)), dim3(BLOCK), 0, as_stream(stream), rows, a_rp, a_ci,
                           a_v, b_rp, b_ci, b_v, pass, cnt, c_rp, c_ci, c_v, overflow);
    else if (method == 2)
        hipLaunchKernelGGL(k_spgemm_wave, dim3(rows), dim3(WAVE), 0, as_stream(stream), rows, a_rp, a_ci, a_v, b_rp,
                           b_ci, b_v, pass, cnt, c_rp, c_ci, c_v, overflow);
    else
        hipLaunchKernelGGL(k_spgemm_rows, dim3(rows), dim3(SPGEMM_THREADS), 0, as_stream(stream), rows, a_rp, a_ci,
                           a_v, b_rp, b_ci, b_v, pass, cnt, c_rp, c_ci, c_v, overflow);
    return launch_status();
}

size_t cfa_scan_counts_ws_bytes(int n) {
    return n < 0 ? 0 : scan_bytes(n);
}

int cfa_scan_counts(int n, const int* cnt, int* out, int* total, void* ws, size_t ws_bytes, cfa_stream_t stream) {
    if (n < 0 || !out || (n > 0 && !cnt)) return CFA_E_ARG;
    size_t need = scan_bytes(n);
    if (need == 0 || !ws || ws_bytes < need) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    const auto in = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0), CountAt{cnt, n});
    const hipError_t e = rocprim::exclusive_scan(ws, need, in, out, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    if (total) hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(WAVE), 0, s, out + n, total);
    return launch_status();
}

int cfa_amg_prolongator(int n, const int* ap_rp, const int* ap_ci, const double* ap_v, const int* agg,
                        const double* p0v, const double* dinv, double omega, int pass, int* cnt, const int* p_rp,
                        int* p_ci, double* p_v, cfa_stream_t stream) {
    if (n < 0 || (pass != 0 && pass != 1)) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    if (!ap_rp || !ap_ci || !ap_v || !agg || !p0v || !dinv) return CFA_E_ARG;
    if ((pass == 0 && !cnt) || (pass == 1 && (!p_rp || !p_ci || !p_v))) return CFA_E_ARG;
    hipLaunchKernelGGL(k_prolongator_rows, dim3(blocks(n)), dim3(BLOCK), 0, as_stream(stream), n, ap_rp, ap_ci, ap_v,
                       agg, p0v, dinv, omega, pass, cnt, p_rp, p_ci, p_v);
    return launch_status();
}

size_t cfa_csr_transpose_ws_bytes(int rows, int cols, int nnz) {
    if (rows < 0 || cols < 0 || nnz < 0) return 0;
    const size_t sort = transpose_sort_bytes(nnz, cols);
    return sort ? align_up(3 * (size_t)nnz * sizeof(int)) + sort : 0;
}

int cfa_csr_transpose(int rows, int cols, int nnz, const int* rp, const int* ci, const double* v, void* ws,
                      size_t ws_bytes, int* t_rp, int* t_ci, double* t_v, cfa_stream_t stream) {
    if (rows < 0 || cols < 0 || nnz < 0 || !t_rp || (rows > 0 && !rp) || (nnz > 0 && (!ci || !v || !t_ci || !t_v)))
        return CFA_E_ARG;
    const size_t need = cfa_csr_transpose_ws_bytes(rows, cols, nnz);
    if (need == 0 || !ws || ws_bytes < need) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    unsigned* keys = static_cast<unsigned*>(ws);
    int* order = reinterpret_cast<int*>(keys + nnz);
    int* row_of = order + nnz;
    void* tmp = static_cast<char*>(ws) + align_up(3 * (size_t)nnz * sizeof(int));
    size_t tmp_bytes = ws_bytes - align_up(3 * (size_t)nnz * sizeof(int));
    if (nnz > 0) {
        hipLaunchKernelGGL(k_row_of_entries, dim3(blocks(rows)), dim3(BLOCK), 0, s, rows, rp, row_of);
        const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, reinterpret_cast<const unsigned*>(ci), keys,
                                                       rocprim::counting_iterator<int>(0), order, (size_t)nnz, 0,
                                                       column_bits(cols), s);
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL(k_transpose_gather, dim3(blocks(nnz)), dim3(BLOCK), 0, s, nnz, order, row_of, v, t_ci, t_v);
    }
    hipLaunchKernelGGL(k_column_starts, dim3(blocks(cols + 1)), dim3(BLOCK), 0, s, cols, nnz, keys, t_rp);
    return launch_status();
}

size_t cfa_spgemm_sorted_ws_bytes(int rows, int cols, int a_nnz, long long products) {
    if (rows < 0 || cols < 0 || a_nnz < 0 || products < 0 || products > 0x7fffffffLL) return 0;
    return sorted_layout(a_nnz, products, column_bits(cols) + column_bits(rows), nullptr, nullptr);
}

int cfa_spgemm_sorted(int rows, int cols, int a_nnz, const int* a_rp, const int* a_ci, const double* a_v,
                      const int* b_rp, const int* b_ci, const double* b_v, int pass, long long products, void* ws,
                      size_t ws_bytes, long long* count, int* c_rp, int* c_ci, double* c_v, cfa_stream_t stream) {
    if (rows < 0 || cols < 0 || a_nnz < 0 || pass < 0 || pass > 2 || products < 0 || !ws) return CFA_E_ARG;
    if ((pass < 2 && !count) || (pass == 2 && !c_rp)) return CFA_E_ARG;
    if (a_nnz > 0 && (!a_rp || !a_ci || !a_v || !b_rp || !b_ci || !b_v)) return CFA_E_ARG;
    const size_t need = cfa_spgemm_sorted_ws_bytes(rows, cols, a_nnz, pass == 0 ? 0 : products);
    if (need == 0 || ws_bytes < need) return CFA_E_ARG;
    const unsigned cb = column_bits(cols), end_bit = cb + column_bits(rows);
    SortedWs w;
    sorted_layout(a_nnz, pass == 0 ? 0 : products, end_bit, ws, &w);
    const hipStream_t s = as_stream(stream);
    const long long P = products;
    if (pass < 2) {
        // products per left entry and their offsets (also the total)
        if (a_nnz > 0)
            hipLaunchKernelGGL(k_entry_products, dim3(blocks(a_nnz)), dim3(BLOCK), 0, s, a_nnz, a_ci, b_rp, w.cnt);
        const auto in = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0), CountAt64{w.cnt, a_nnz});
        size_t tb = w.tmp_bytes;
        hipError_t e = rocprim::exclusive_scan(w.tmp, tb, in, w.off, 0LL, (size_t)a_nnz + 1,
                                               rocprim::plus<long long>(), s);
        if (e != hipSuccess) return (int)e;
        if (pass == 0) {
            e = hipMemcpyAsync(count, w.off + a_nnz, sizeof(long long), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            return e == hipSuccess ? launch_status() : (int)e;
        }
        // list, sort, mark the groups; *count = groups
        if (P > 0) {
            hipLaunchKernelGGL(k_row_of_entries, dim3(blocks(rows)), dim3(BLOCK), 0, s, rows, a_rp, w.row_of);
            hipLaunchKernelGGL(k_expand_products, dim3(blocks(a_nnz)), dim3(BLOCK), 0, s, a_nnz, w.row_of, a_ci, a_v,
                               b_rp, b_ci, b_v, w.off, cb, w.key[0], w.val[0]);
            tb = w.tmp_bytes;
            e = rocprim::radix_sort_pairs(w.tmp, tb, w.key[0], w.key[1], w.val[0], w.val[1], (size_t)P, 0, end_bit, s);
            if (e != hipSuccess) return (int)e;
            hipLaunchKernelGGL(k_group_heads, dim3(blocks_for(P)), dim3(BLOCK), 0, s, P, w.key[1], w.head);
        }
        const auto inh = rocprim::make_transform_iterator(rocprim::counting_iterator<int>(0), CountAt{w.head, (int)P});
        tb = w.tmp_bytes;
        e = rocprim::exclusive_scan(w.tmp, tb, inh, w.gidx, 0, (size_t)P + 1, rocprim::plus<int>(), s);
        if (e != hipSuccess) return (int)e;
        int groups = 0;
        e = hipMemcpyAsync(&groups, w.gidx + P, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return (int)e;
        *count = groups;
        return launch_status();
    }
    // pass 2: the sums, columns and row starts (the workspace of pass 1, unchanged)
    int groups = 0;
    if (P > 0) {
        hipError_t e = hipMemcpyAsync(&groups, w.gidx + P, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return (int)e;
        if (groups > 0 && (!c_ci || !c_v)) return CFA_E_ARG;
        hipLaunchKernelGGL(k_group_sums, dim3(blocks_for(P)), dim3(BLOCK), 0, s, P, w.key[1], w.val[1], w.gidx, cb, c_ci,
                           c_v, w.head);
    }
    hipLaunchKernelGGL(k_row_starts, dim3(blocks(rows + 1)), dim3(BLOCK), 0, s, rows, groups, w.head, c_rp);
    return launch_status();
}

}  // extern "C"
