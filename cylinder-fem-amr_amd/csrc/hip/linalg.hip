// Sparse linear algebra for gfx950: CSR SpMV (several lanes per row, wave64 reductions), dot
// products with fixed-order two-stage reductions, multicolour ILU(0), the SA-AMG V-cycle and the
// Krylov solvers of src/cfa/linalg.cuh.
//
// Krylov design: the reference copies every dot product to the host (three blocking transfers
// per CG iteration, linalg.cuh:172-177). Here all scalars stay on the device: a one-workgroup
// "final" kernel reduces the block partials in a fixed order and applies the scalar logic of the
// reference (coefficients, breakdown guards, stopping rules), and the iteration counter lives on
// the device too. Every vector kernel reads a device state word and returns at once when the
// solve has stopped. One iteration is therefore the same kernel sequence every time: it is
// captured once as a HIP graph and replayed (about 1.7 us per kernel node on MI355X against
// 2.5 us per stream launch), with the host a few iterations ahead, following a progress word
// that the final kernels write to mapped pinned memory, so the queue never drains.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "cfa_device.hpp"

using namespace cfa_dev;

namespace {

constexpr int RED_BLOCKS = 1024;   // partial sums per reduction (grid-stride kernels)
constexpr int OP_BLOCKS = 4096;    // grid cap of the fused operator kernels (partials <= workspace)
constexpr int LOOKAHEAD = 3;       // iterations enqueued ahead of the last one known to have finished
constexpr int GRAPH_CACHE = 4;     // captured iteration graphs kept per solver kind

// Solve state (device ints after the scalar area of the workspace)
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_STOPPED = 2, ST_TRIVIAL = 3 };

__device__ __forceinline__ bool halted(const int* state) { return state && state[0] != ST_RUNNING; }

// Loads a row's range [k0, k1) and the solve state together and reports whether the solve runs:
// both loads are in flight at once, so the state word does not add a round trip in front of the
// row's loads (measured: 0.25-0.3 us of a 3-4.4 us sweep on 18k-115k rows). The empty asm
// statement makes the three values complete before the early exit.
// Rows past the end get the empty range [ptr[rows], ptr[rows]); the loads are unconditional (and
// in bounds: ptr has rows + 1 entries) so that no branch separates them.
__device__ __forceinline__ bool running_with_range(const int* __restrict__ ptr, int row, int rows,
                                                   const int* __restrict__ state, int& k0, int& k1) {
    const int r = row < rows ? row : rows;
    k0 = ptr[r];
    k1 = ptr[r + (row < rows ? 1 : 0)];
    const int sv = state ? state[0] : ST_RUNNING;
    asm volatile("" ::"v"(k0), "v"(k1), "s"(sv));
    return sv == ST_RUNNING;
}

__device__ __forceinline__ void publish(int* status, int progress, int state) {
    if (!status) return;
    __hip_atomic_store(&status[1], state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&status[0], progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- CSR products --------------------------------------------------------------------------------
enum CsrOp { OP_SPMV, OP_ADD, OP_RESID, OP_JACOBI };

// y = A x | y += A x | y = b - A x | y = x + omega dinv (b - A x)   (src/cfa/linalg.cuh:884-901).
// T = double, or float for the mixed-precision cycle (values, vectors and arithmetic in FP32).
template <int TPR, int OP, typename T = double>
__global__ void __launch_bounds__(BLOCK)
k_csr_op(int rows, const int* __restrict__ rp, const int* __restrict__ ci, const T* __restrict__ val,
         const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ dinv, T omega, T* __restrict__ y,
         const int* __restrict__ state) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = g / TPR, lane = g % TPR;
    int k0, k1;
    if (!running_with_range(rp, row, rows, state, k0, k1)) return;
    T s = 0;
    row_entries<TPR>(ci, val, k0, k1, lane, [&](T a, int j) { s += a * x[j]; });
    s = group_sum<TPR>(s);
    if (row >= rows || lane != 0) return;
    if (OP == OP_SPMV) y[row] = s;
    else if (OP == OP_ADD) y[row] += s;
    else if (OP == OP_RESID) y[row] = b[row] - s;
    else y[row] = x[row] + omega * dinv[row] * (b[row] - s);
}

// First two steps of the V-cycle smoother fused: x0 = omega dinv b (src/cfa/linalg.cuh:864-868),
// y = x0 + omega dinv (b - A x0), with x0 formed on the fly from b and dinv at the neighbours.
template <int TPR, typename T = double>
__global__ void __launch_bounds__(BLOCK)
k_jacobi_from_zero(int rows, const int* __restrict__ rp, const int* __restrict__ ci, const T* __restrict__ val,
                   const T* __restrict__ b, const T* __restrict__ dinv, T omega, T* __restrict__ y,
                   const int* __restrict__ state) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = g / TPR, lane = g % TPR;
    int k0, k1;
    if (!running_with_range(rp, row, rows, state, k0, k1)) return;
    T s = 0;
    row_entries<TPR>(ci, val, k0, k1, lane, [&](T a, int j) { s += a * (omega * dinv[j] * b[j]); });
    s = group_sum<TPR>(s);
    if (row >= rows || lane != 0) return;
    const T x0 = omega * dinv[row] * b[row];
    y[row] = x0 + omega * dinv[row] * (b[row] - s);
}

template <typename T = double>
__global__ void k_jacobi_init(int n, T omega, const T* __restrict__ dinv, const T* __restrict__ b, T* __restrict__ x,
                              const int* __restrict__ state) {
    if (halted(state)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = omega * dinv[i] * b[i];
}

// y = Ainv x with the dense coarse inverse, one wavefront per row (src/cfa/linalg.cuh:929-937);
// FP64 arithmetic also when x and y are FP32 (the coarse solve stays FP64 in the mixed cycle)
template <typename T = double>
__global__ void __launch_bounds__(BLOCK)
k_dense_matvec(int n, const double* __restrict__ A, const T* __restrict__ x, T* __restrict__ y,
               const int* __restrict__ state) {
    if (halted(state)) return;
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, lane = threadIdx.x & (WAVE - 1);
    double s = 0.0;
    if (i < n)
        for (int j = lane; j < n; j += WAVE) s += A[(size_t)i * n + j] * (double)x[j];
    s = wave_sum(s);
    if (i < n && lane == 0) y[i] = (T)s;
}

// Level data of the double or the float cycle
template <typename T> struct Lv;
template <> struct Lv<double> {
    __host__ __device__ static const double* a(const cfa_amg_level_t& L) { return L.a_v; }
    __host__ __device__ static const double* p(const cfa_amg_level_t& L) { return L.p_v; }
    __host__ __device__ static const double* r(const cfa_amg_level_t& L) { return L.r_v; }
    __host__ __device__ static const double* dinv(const cfa_amg_level_t& L) { return L.dinv; }
    __host__ __device__ static const double* sv(const cfa_amg_level_t& L) { return L.s_v; }
    __host__ __device__ static double* x(const cfa_amg_level_t& L) { return L.x; }
    __host__ __device__ static double* b(const cfa_amg_level_t& L) { return L.b; }
    __host__ __device__ static double* res(const cfa_amg_level_t& L) { return L.r; }
    __host__ __device__ static double* ax(const cfa_amg_level_t& L) { return L.ax; }
    static double* cb(const cfa_amg_t* a) { return a->cb; }
    static double* cx(const cfa_amg_t* a) { return a->cx; }
};
template <> struct Lv<float> {
    __host__ __device__ static const float* a(const cfa_amg_level_t& L) { return L.a_v32; }
    __host__ __device__ static const float* p(const cfa_amg_level_t& L) { return L.p_v32; }
    __host__ __device__ static const float* r(const cfa_amg_level_t& L) { return L.r_v32; }
    __host__ __device__ static const float* dinv(const cfa_amg_level_t& L) { return L.dinv32; }
    __host__ __device__ static const float* sv(const cfa_amg_level_t& L) { return L.s_v32; }
    __host__ __device__ static float* x(const cfa_amg_level_t& L) { return L.x32; }
    __host__ __device__ static float* b(const cfa_amg_level_t& L) { return L.b32; }
    __host__ __device__ static float* res(const cfa_amg_level_t& L) { return L.r32; }
    __host__ __device__ static float* ax(const cfa_amg_level_t& L) { return L.ax32; }
    static float* cb(const cfa_amg_t* a) { return a->cb32; }
    static float* cx(const cfa_amg_t* a) { return a->cx32; }
};

// ---- reductions ------------------------------------------------------------------------------------
// partial[b] = sum over the block's grid-stride share of x . y
__global__ void __launch_bounds__(BLOCK)
k_dot_partial(int n, const double* __restrict__ x, const double* __restrict__ y, double* __restrict__ part,
              const int* __restrict__ state) {
    __shared__ double lds[BLOCK / WAVE];
    if (halted(state)) return;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += x[i] * y[i];
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ void k_sum_to(int nb, const double* __restrict__ part, double* __restrict__ out, int take_sqrt) {
    __shared__ double lds[BLOCK / WAVE];
    const double s = sum_partials(part, nb, 1, lds);
    if (threadIdx.x == 0) *out = take_sqrt ? sqrt(s) : s;
}

inline int red_blocks(long long n) {
    int nb = blocks_for(n > 0 ? n : 1);
    return nb > RED_BLOCKS ? RED_BLOCKS : nb;
}

// ---- elementwise ------------------------------------------------------------------------------------
__global__ void k_axpy(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void k_xpay(int n, const double* __restrict__ x, double b, double* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] + b * y[i];
}
__global__ void k_vec_mul(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out,
                          const int* __restrict__ state) {
    if (halted(state)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * b[i];
}
__global__ void k_fill_zero(int n, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.0;
}
__global__ void k_copy(int n, const double* __restrict__ a, double* __restrict__ out, const int* __restrict__ state) {
    if (halted(state)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i];
}
// inverse diagonal, 1 where |d| <= eps (src/cfa/linalg.cuh:28-48 with eps 1e-14, :871-878 with
// eps 1e-300; a missing diagonal counts as 1 resp. 0, both giving 1)
__global__ void k_inv_diag(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ val, double eps, double* __restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = eps > 1.0e-100 ? 1.0 : 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) { d = val[k]; break; }
    inv[i] = (fabs(d) > eps) ? (1.0 / d) : 1.0;
}

template <int TPR, int NRHS>
__global__ void __launch_bounds__(BLOCK)
k_spmv_multi(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ val,
             const double* __restrict__ x, double* __restrict__ y) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = g / TPR, lane = g % TPR;
    double s[NRHS];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) s[c] = 0.0;
    if (row < n)
        row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane, [&](double a, int j) {
#pragma unroll
            for (int c = 0; c < NRHS; ++c) s[c] += a * x[(size_t)c * n + j];
        });
#pragma unroll
    for (int c = 0; c < NRHS; ++c) s[c] = group_sum<TPR>(s[c]);
    if (row < n && lane == 0)
#pragma unroll
        for (int c = 0; c < NRHS; ++c) y[(size_t)c * n + row] = s[c];
}

// ======================================================================================================
// PCG (src/cfa/linalg.cuh:183-256)
// ======================================================================================================
enum { S_BNORM = 0, S_RZ_OLD, S_PQ, S_ALPHA, S_RNORM, S_REL, S_BETA, S_TOL };
// ints: [0] state, [1] final iteration count, [2] current iteration

// ||b||; x = 0 and stop when ||b|| < 1e-14 (linalg.cuh:205-209)
__global__ void k_pcg_init_bnorm(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                 int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[BLOCK / WAVE];
    const double bb = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    const double b_norm = sqrt(bb);
    sc[S_BNORM] = b_norm;
    st[1] = 0;
    st[2] = 0;
    st[0] = b_norm < 1.0e-14 ? ST_TRIVIAL : ST_RUNNING;
    publish(status, 0, st[0]);
}

__global__ void k_zero_if_trivial(int n, double* __restrict__ x, const int* __restrict__ st) {
    if (st[0] != ST_TRIVIAL) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// r = b - q; partial r.r
__global__ void __launch_bounds__(BLOCK)
k_pcg_residual(int n, const double* __restrict__ b, const double* __restrict__ q, double* __restrict__ r,
               double* __restrict__ part, const int* __restrict__ st) {
    __shared__ double lds[BLOCK / WAVE];
    if (halted(st)) return;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ri = b[i] + (-1.0) * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// initial residual test (linalg.cuh:217-218): converged with 0 iterations
__global__ void k_pcg_init_rnorm(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                 int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[BLOCK / WAVE];
    if (st[0] != ST_RUNNING) return;
    const double rr = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    const double r_norm = sqrt(rr);
    if (r_norm < sc[S_TOL] * sc[S_BNORM]) {
        st[0] = ST_CONVERGED;
        st[1] = 0;
        publish(status, 0, ST_CONVERGED);
    }
}

// alpha = rz_old / (p.q); breakdown when |p.q| < 1e-30
__global__ void k_pcg_final_alpha(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                  int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[BLOCK / WAVE];
    if (st[0] != ST_RUNNING) return;
    const double pq = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    sc[S_PQ] = pq;
    if (fabs(pq) < 1.0e-30) {
        st[0] = ST_STOPPED;
        st[1] = st[2];
        publish(status, st[2], ST_STOPPED);
        return;
    }
    sc[S_ALPHA] = sc[S_RZ_OLD] / pq;
}

// x += alpha p; r -= alpha q; partial r.r
__global__ void __launch_bounds__(BLOCK)
k_pcg_update_xr(int n, const double* __restrict__ sc, const double* __restrict__ p, const double* __restrict__ q,
                double* __restrict__ x, double* __restrict__ r, double* __restrict__ part, const int* __restrict__ st) {
    __shared__ double lds[BLOCK / WAVE];
    if (halted(st)) return;
    const double alpha = sc[S_ALPHA];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] + (-alpha) * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// ||r|| < tol ||b|| -> converged after it + 1 iterations
__global__ void k_pcg_final_rnorm(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                  int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[BLOCK / WAVE];
    if (st[0] != ST_RUNNING) return;
    const double rr = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    const double r_norm = sqrt(rr);
    sc[S_RNORM] = r_norm;
    sc[S_REL] = r_norm / sc[S_BNORM];
    if (r_norm < sc[S_TOL] * sc[S_BNORM]) {
        st[0] = ST_CONVERGED;
        st[1] = st[2] + 1;
        publish(status, st[2] + 1, ST_CONVERGED);
    }
}

// beta = rz_new / (rz_old + 1e-30); rz_old = rz_new; next iteration (first: prologue, rz only)
__global__ void k_pcg_final_beta(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                 int* __restrict__ st, int* __restrict__ status, int first) {
    __shared__ double lds[BLOCK / WAVE];
    if (st[0] != ST_RUNNING) return;
    const double rz = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    if (!first) {
        sc[S_BETA] = rz / (sc[S_RZ_OLD] + 1.0e-30);
        st[2] += 1;
    }
    sc[S_RZ_OLD] = rz;
    publish(status, st[2], ST_RUNNING);
}

// p = z + beta p
__global__ void k_pcg_p(int n, const double* __restrict__ z, const double* __restrict__ sc, double* __restrict__ p,
                        const int* __restrict__ st) {
    if (halted(st)) return;
    const double beta = sc[S_BETA];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// Fused operator, stage 1: q = Kp p (CSR) and g = M^-1 B p, both on the same row pattern
template <int TPR>
__global__ void __launch_bounds__(BLOCK)
k_apply_K_B(int n, int nnz, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ val,
            const double* __restrict__ oss, const double* __restrict__ inv_m, const double* __restrict__ in,
            double* __restrict__ q, double* __restrict__ gx, double* __restrict__ gy, const int* __restrict__ st) {
    if (halted(st)) return;
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int lane = threadIdx.x % TPR;
    const RowValues<3, double> kb{{val, oss, oss + (size_t)nnz}};
    for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / TPR; row < n; row += groups) {
        double s = 0.0, sx = 0.0, sy = 0.0;
        row_entries_n<TPR>(ci, kb, rp[row], rp[row + 1], lane, [&](const double* v, int j) {
            const double xj = in[j];
            s += v[0] * xj;
            sx += v[1] * xj;
            sy += v[2] * xj;
        });
        s = group_sum<TPR>(s);
        sx = group_sum<TPR>(sx);
        sy = group_sum<TPR>(sy);
        if (lane == 0) {
            q[row] = s;
            gx[row] = sx * inv_m[row];
            gy[row] = sy * inv_m[row];
        }
    }
}

// Fused operator, stage 2: q -= B^T g on non-Dirichlet rows; partial p.q per block. The grid is
// capped (grid-stride over rows) so the partials fit the workspace.
template <int TPR>
__global__ void __launch_bounds__(BLOCK)
k_apply_BT_dot(int n, int nnz, const int* __restrict__ rp, const int* __restrict__ ci,
               const double* __restrict__ oss, const int* __restrict__ tag, OutletMode om,
               const double* __restrict__ gx, const double* __restrict__ gy, const double* __restrict__ p,
               double* __restrict__ q, double* __restrict__ part, const int* __restrict__ st) {
    __shared__ double lds[BLOCK / WAVE];
    if (halted(st)) return;
    const double* BTx = oss + 2 * (size_t)nnz;
    const double* BTy = oss + 3 * (size_t)nnz;
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int lane = threadIdx.x % TPR;
    const RowValues<2, double> bt{{BTx, BTy}};
    double contrib = 0.0;
    for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / TPR; row < n; row += groups) {
        const bool active = !pressure_dirichlet(row, tag[row], om);
        double s = 0.0;
        if (active)
            row_entries_n<TPR>(ci, bt, rp[row], rp[row + 1], lane,
                               [&](const double* v, int c) { s += v[0] * gx[c] + v[1] * gy[c]; });
        s = group_sum<TPR>(s);
        if (lane == 0) {
            double qi = q[row];
            if (active) {
                qi -= s;
                q[row] = qi;
            }
            contrib += p[row] * qi;
        }
    }
    contrib = block_sum(contrib, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = contrib;
}

// ======================================================================================================
// PCG with the scalar reductions folded into the kernels that consume them
// ======================================================================================================
// A kernel that needs a scalar (beta, alpha, the residual test) sums its producer's block
// partials itself, in the same fixed order in every block, so all blocks reach the same value
// and the same decision without a separate one-workgroup kernel; only block 0 writes the state
// words. The search direction alternates between two buffers so that the operator kernel can
// form p_k = z_k + beta_{k-1} p_{k-1} on the fly while it applies the operator (the reference's
// separate k_xpay, linalg.cuh:252). r.z partials alternate between two buffers as well: the
// operator of iteration k reads r_k.z_k and r_{k-1}.z_{k-1}, while the preconditioner of the same
// iteration writes r_{k+1}.z_{k+1} over the older one.
enum { ST_CAPPED = 4 };

struct PcgArgs {
    int n, nnz, max_iter;
    const int* rp; const int* ci; const double* val;
    const double* oss; const double* inv_m; const int* tag; double* gx; double* gy;
    OutletMode om;
    double* x; double* r; double* z; double* q; double* p[2];
    double* prz[2]; double* ppq; double* prr;
    int nb_rz, nb_rr;
    const double* inv_diag;
    double* sc; int* st; int* status;
};

// state word read once per block (block-uniform early exits in kernels with barriers)
__device__ __forceinline__ int block_state(const int* st, int* it_out) {
    __shared__ int s_state, s_it;
    if (threadIdx.x == 0) { s_state = st[0]; s_it = st[2]; }
    __syncthreads();
    if (it_out) *it_out = s_it;
    return s_state;
}

// Operator of iteration k (parity h): p_k formed on the fly and stored to p[h]; q = Kp p_k, and
// with OSS g = M^-1 B p_k for the second stage, else the p.q partials here. Iteration cap check.
template <int TPR, bool OSS>
__global__ void __launch_bounds__(BLOCK)
k_fp_operator(PcgArgs a, int h) {
    __shared__ double lds[2 * BLOCK / WAVE];
    int it;
    if (block_state(a.st, &it) != ST_RUNNING) return;
    if (it >= a.max_iter) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { a.st[0] = ST_CAPPED; publish(a.status, it, ST_CAPPED); }
        return;
    }
    const bool first = it == 0;
    double beta = 0.0;
    if (!first) {
        double rz[2];
        sum_partials_k<2>({a.prz[h], a.prz[1 - h]}, {a.nb_rz, a.nb_rz}, 1, rz, lds);
        beta = rz[0] / (rz[1] + 1.0e-30);
    }
    const double* __restrict__ z = a.z;
    const double* __restrict__ po = a.p[1 - h];
    double* __restrict__ pn = a.p[h];
    const int* __restrict__ rp = a.rp;
    const int* __restrict__ ci = a.ci;
    const double* __restrict__ val = a.val;
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int lane = threadIdx.x % TPR;
    const RowValues<3, double> kb{{val, a.oss, a.oss + (size_t)a.nnz}};
    double contrib = 0.0;
    for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / TPR; row < a.n; row += groups) {
        double s = 0.0, sx = 0.0, sy = 0.0;
        if (!OSS) {
            row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane, [&](double av, int j) {
                s += av * (first ? z[j] : fma(beta, po[j], z[j]));
            });
        } else {
            row_entries_n<TPR>(ci, kb, rp[row], rp[row + 1], lane, [&](const double* v, int j) {
                const double pj = first ? z[j] : fma(beta, po[j], z[j]);
                s += v[0] * pj;
                sx += v[1] * pj;
                sy += v[2] * pj;
            });
        }
        s = group_sum<TPR>(s);
        if (OSS) {
            sx = group_sum<TPR>(sx);
            sy = group_sum<TPR>(sy);
        }
        if (lane == 0) {
            const double pi = first ? z[row] : fma(beta, po[row], z[row]);
            pn[row] = pi;
            a.q[row] = s;
            if (OSS) {
                a.gx[row] = sx * a.inv_m[row];
                a.gy[row] = sy * a.inv_m[row];
            } else {
                contrib += pi * s;
            }
        }
    }
    if (!OSS) {
        contrib = block_sum(contrib, lds);
        if (threadIdx.x == 0) a.ppq[blockIdx.x] = contrib;
    }
}

// alpha = r_k.z_k / (p_k.q) with the breakdown guard, x += alpha p, r -= alpha q, r.r partials;
// the completed-iteration counter advances here.
__global__ void __launch_bounds__(BLOCK)
k_fp_update(PcgArgs a, int h, int nb_pq) {
    __shared__ double lds[2 * BLOCK / WAVE];
    int it;
    if (block_state(a.st, &it) != ST_RUNNING) return;
    double sums[2];   // p.q, r.z
    sum_partials_k<2>({a.ppq, a.prz[h]}, {nb_pq, a.nb_rz}, 1, sums, lds);
    const double pq = sums[0];
    if (fabs(pq) < 1.0e-30) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.st[0] = ST_STOPPED;
            a.st[1] = it;
            publish(a.status, it, ST_STOPPED);
        }
        return;
    }
    const double alpha = sums[1] / pq;
    const double* __restrict__ p = a.p[h];
    const double* __restrict__ q = a.q;
    double* __restrict__ x = a.x;
    double* __restrict__ r = a.r;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] + (-alpha) * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) {
        a.prr[blockIdx.x] = acc;
        if (blockIdx.x == 0) a.st[2] = it + 1;
    }
}

// Residual test of the iteration (||r|| < tol ||b||, linalg.cuh:241-247) from the r.r
// partials; returns true (in every thread of every block) when the solve goes on.
__device__ __forceinline__ bool fp_residual_test(const double* prr, int nb_rr, double* sc, int* st, int* status,
                                                 double* lds) {
    const double r_norm = sqrt(sum_partials(prr, nb_rr, 1, lds));
    const bool done = r_norm < sc[S_TOL] * sc[S_BNORM];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc[S_RNORM] = r_norm;
        sc[S_REL] = r_norm / sc[S_BNORM];
        const int it = st[2];
        if (done) {
            st[0] = ST_CONVERGED;
            st[1] = it;
        }
        publish(status, it, done ? ST_CONVERGED : ST_RUNNING);
    }
    return !done;
}

// One-workgroup residual test, for preconditioners whose first kernel cannot take it
__global__ void k_fp_check(const double* prr, int nb_rr, double* sc, int* st, int* status) {
    __shared__ double lds[BLOCK / WAVE];
    if (block_state(st, nullptr) != ST_RUNNING) return;
    (void)fp_residual_test(prr, nb_rr, sc, st, status, lds);
}

// First two V-cycle sweeps from x = 0 on level 0 (as k_jacobi_from_zero), preceded by the
// residual test of the iteration.
template <int TPR>
__global__ void __launch_bounds__(BLOCK)
k_jacobi_from_zero_check(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                         const double* __restrict__ val, const double* __restrict__ b,
                         const double* __restrict__ dinv, double omega, double* __restrict__ y,
                         const double* __restrict__ prr, int nb_rr, double* sc, int* st, int* status) {
    __shared__ double lds[BLOCK / WAVE];
    if (block_state(st, nullptr) != ST_RUNNING) return;
    if (!fp_residual_test(prr, nb_rr, sc, st, status, lds)) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = g / TPR, lane = g % TPR;
    double s = 0.0;
    if (row < rows)
        row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane,
                         [&](double a, int j) { s += a * (omega * dinv[j] * b[j]); });
    s = group_sum<TPR>(s);
    if (row >= rows || lane != 0) return;
    const double x0 = omega * dinv[row] * b[row];
    y[row] = x0 + omega * dinv[row] * (b[row] - s);
}

// Last V-cycle sweep on level 0: y = x + omega dinv (b - A x) written as FP64 (the
// preconditioned residual z), with the r.z block partials when part is set (rdot = the FP64
// residual r; b itself in the FP64 cycle). Grid-stride, capped grid.
template <int TPR, typename T = double>
__global__ void __launch_bounds__(BLOCK)
k_jacobi_dot(int rows, const int* __restrict__ rp, const int* __restrict__ ci, const T* __restrict__ val,
             const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ dinv, T omega,
             double* __restrict__ y, const double* __restrict__ rdot, double* __restrict__ part,
             const int* __restrict__ st) {
    __shared__ double lds[BLOCK / WAVE];
    if (st && block_state(st, nullptr) != ST_RUNNING) return;
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int lane = threadIdx.x % TPR;
    double contrib = 0.0;
    for (int row = (blockIdx.x * blockDim.x + threadIdx.x) / TPR; row < rows; row += groups) {
        T s = 0;
        row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane, [&](T a, int j) { s += a * x[j]; });
        s = group_sum<TPR>(s);
        if (lane == 0) {
            const double yi = (double)(x[row] + omega * dinv[row] * (b[row] - s));
            y[row] = yi;
            if (part) contrib += rdot[row] * yi;
        }
    }
    if (part) {
        contrib = block_sum(contrib, lds);
        if (threadIdx.x == 0) part[blockIdx.x] = contrib;
    }
}

// ---- Sliced-ELL rows (the V-cycle's level-0 sweeps) --------------------------------------------
// The operator's rows in slices of 64 positions, one wavefront each, stored column-major inside a
// slice: entry k of the row at position p = 64 q + l sits at sp[q] + 64 k + l, a slice holds
// (sp[q + 1] - sp[q]) / 64 slots per position, and srow[p] names the row (-1 past the last row).
// Inside each window of SELL_WINDOW positions the rows are ordered by decreasing length (ties by
// row index), so the rows of a slice have nearly the same length and little is padded; padded
// slots hold column 0 and value 0. Each entry load of a wavefront is one contiguous run (512 bytes
// in FP64), where the CSR kernels' loads scatter over the rows of a lane group: on a 7.35 M-row
// matrix of the uniform meshes' 5- and 9-entry rows a Jacobi sweep takes 175 us against 257 us.
// One lane per row; the row's products are summed in the order of the CSR kernels with G lanes per
// row (G interleaved partial sums, joined by the lane group's tree), so the results are the CSR
// kernels' bit for bit. Each cache line of the copy is read by one load instruction, so a sweep
// whose bytes exceed the 256 MB Infinity Cache loads it non-temporally, leaving the caches to the
// gathered vector (uniform level 7: 188 -> 160-166 us); a smaller one keeps plain loads, so that
// the copy stays cached between sweeps (on the level-6 adaptive meshes the hint cost 8.0 -> 9.2 us).
constexpr int SELL_SLICE = 64;
constexpr int SELL_WINDOW = 256;
constexpr double SELL_STREAM_BYTES = 256.0 * 1024 * 1024;

__host__ __device__ inline int sell_positions(int rows) { return (rows + SELL_SLICE - 1) / SELL_SLICE * SELL_SLICE; }

// The position's row, the base of its slots and the slice's length; false when `state` says the
// solve has stopped (its word loaded together with the slice's offsets, as in running_with_range).
__device__ __forceinline__ bool sell_slot(const int* __restrict__ sp, const int* __restrict__ srow, int p,
                                          const int* __restrict__ state, int& row, int& base, int& len) {
    const int q = p / SELL_SLICE;
    const int k0 = sp[q], k1 = sp[q + 1];
    row = srow[p];
    const int sv = state ? state[0] : ST_RUNNING;
    asm volatile("" ::"v"(k0), "v"(k1), "v"(row), "s"(sv));
    base = k0 + p % SELL_SLICE;
    len = (k1 - k0) / SELL_SLICE;
    return sv == ST_RUNNING;
}

// sum over the row's entries k of a_k xv(j_k), every index and value loaded before any is used
// (slots past the slice's length give (0, column 0) without a load, so the gathers are
// unconditional and add exact zeros), in the order of row_entries<G> + group_sum<G>: entry k goes
// to partial k mod G, each partial in entry order, the partials joined as the lane tree joins them
// (G = 2: p0 + p1; G = 4: (p0 + p2) + (p1 + p3)). NT: non-temporal loads of the copy.
template <int MAXL, int G, bool NT, typename T, typename F>
__device__ __forceinline__ T sell_row_sum(const int* __restrict__ ci, const T* __restrict__ val, int base, int len,
                                         F xv) {
    int c[MAXL];
    T v[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        const bool ok = k < len;
        const int at = base + SELL_SLICE * k;
        if constexpr (NT) {
            c[k] = ok ? __builtin_nontemporal_load(ci + at) : 0;
            v[k] = ok ? __builtin_nontemporal_load(val + at) : T(0);
        } else {
            c[k] = ok ? ci[at] : 0;
            v[k] = ok ? val[at] : T(0);
        }
    }
    T acc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) acc[g] = 0;
#pragma unroll
    for (int k = 0; k < MAXL; ++k) acc[k % G] += v[k] * xv(c[k]);
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
        for (int g = 0; g < o; ++g) acc[g] += acc[g + o];
    return acc[0];
}

// y = b - A x (OP_RESID) | y = x + omega dinv (b - A x) (OP_JACOBI) on the sliced-ELL copy
template <int MAXL, int G, bool NT, int OP, typename T>
__global__ void __launch_bounds__(BLOCK)
k_sell_op(int positions, const int* __restrict__ sp, const int* __restrict__ srow, const int* __restrict__ ci,
          const T* __restrict__ val, const T* __restrict__ x, const T* __restrict__ b, const T* __restrict__ dinv,
          T omega, T* __restrict__ y, const int* __restrict__ state) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= positions) return;
    int row, base, len;
    if (!sell_slot(sp, srow, p, state, row, base, len)) return;
    const T s = sell_row_sum<MAXL, G, NT>(ci, val, base, len, [&](int j) { return x[j]; });
    if (row < 0) return;
    if (OP == OP_RESID) y[row] = b[row] - s;
    else y[row] = x[row] + omega * dinv[row] * (b[row] - s);
}

// k_jacobi_from_zero on the sliced-ELL copy
template <int MAXL, int G, bool NT, typename T>
__global__ void __launch_bounds__(BLOCK)
k_sell_from_zero(int positions, const int* __restrict__ sp, const int* __restrict__ srow, const int* __restrict__ ci,
                 const T* __restrict__ val, const T* __restrict__ b, const T* __restrict__ dinv, T omega,
                 T* __restrict__ y, const int* __restrict__ state) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= positions) return;
    int row, base, len;
    if (!sell_slot(sp, srow, p, state, row, base, len)) return;
    const T s = sell_row_sum<MAXL, G, NT>(ci, val, base, len, [&](int j) { return omega * dinv[j] * b[j]; });
    if (row < 0) return;
    const T x0 = omega * dinv[row] * b[row];
    y[row] = x0 + omega * dinv[row] * (b[row] - s);
}

// k_jacobi_from_zero_check on the sliced-ELL copy: the iteration's residual test, then the sweeps
template <int MAXL, int G, bool NT>
__global__ void __launch_bounds__(BLOCK)
k_sell_from_zero_check(int positions, const int* __restrict__ sp, const int* __restrict__ srow,
                       const int* __restrict__ ci, const double* __restrict__ val, const double* __restrict__ b,
                       const double* __restrict__ dinv, double omega, double* __restrict__ y,
                       const double* __restrict__ prr, int nb_rr, double* sc, int* st, int* status) {
    __shared__ double lds[BLOCK / WAVE];
    if (block_state(st, nullptr) != ST_RUNNING) return;
    if (!fp_residual_test(prr, nb_rr, sc, st, status, lds)) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= positions) return;
    int row, base, len;
    (void)sell_slot(sp, srow, p, nullptr, row, base, len);
    const double s = sell_row_sum<MAXL, G, NT>(ci, val, base, len, [&](int j) { return omega * dinv[j] * b[j]; });
    if (row < 0) return;
    const double x0 = omega * dinv[row] * b[row];
    y[row] = x0 + omega * dinv[row] * (b[row] - s);
}

// The positions of one window's rows (a workgroup per window of SELL_WINDOW rows): each row's rank
// among the window's rows by decreasing length, ties by row index; the slices' entry counts (64
// times the length of their first, longest row) and the longest row (atomic max into info[1]).
__global__ void __launch_bounds__(SELL_WINDOW)
k_sell_plan(int rows, int positions, const int* __restrict__ rp, int* __restrict__ srow, int* __restrict__ cnt,
            int* __restrict__ info) {
    __shared__ int len_of[SELL_WINDOW];
    __shared__ int len_at[SELL_WINDOW];
    const int t = threadIdx.x, w0 = blockIdx.x * SELL_WINDOW, r = w0 + t;
    const int L = r < rows ? rp[r + 1] - rp[r] : -1;
    len_of[t] = L;
    __syncthreads();
    int rank = 0;
    for (int u = 0; u < SELL_WINDOW; ++u) {
        const int Lu = len_of[u];
        rank += (Lu > L || (Lu == L && u < t)) ? 1 : 0;
    }
    len_at[rank] = L;
    if (w0 + rank < positions) srow[w0 + rank] = L >= 0 ? r : -1;
    if (L > 0) atomicMax(&info[1], L);
    __syncthreads();
    if (t < SELL_WINDOW / SELL_SLICE) {
        const int q = w0 / SELL_SLICE + t;
        if (q * SELL_SLICE < positions) cnt[q] = SELL_SLICE * max(len_at[t * SELL_SLICE], 0);
    }
}

// The slots of each position: its row's entries in CSR order, then (column 0, value 0) up to the
// slice's length
__global__ void k_sell_fill(int positions, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ val, const int* __restrict__ srow, const int* __restrict__ sp,
                            int* __restrict__ sci, double* __restrict__ sv) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= positions) return;
    const int q = p / SELL_SLICE;
    const int base = sp[q] + p % SELL_SLICE, len = (sp[q + 1] - sp[q]) / SELL_SLICE;
    const int row = srow[p];
    const int k0 = row >= 0 ? rp[row] : 0, rl = row >= 0 ? rp[row + 1] - k0 : 0;
    for (int k = 0; k < len; ++k) {
        const bool ok = k < rl;
        sci[base + SELL_SLICE * k] = ok ? ci[k0 + k] : 0;
        sv[base + SELL_SLICE * k] = ok ? val[k0 + k] : 0.0;
    }
}

// Entry of the FP32 cycle: b32 = (float) r, after the iteration's residual test when prr is set
__global__ void k_to_f32_check(int n, const double* __restrict__ r, float* __restrict__ b32,
                               const double* __restrict__ prr, int nb_rr, double* sc, int* st, int* status) {
    __shared__ double lds[BLOCK / WAVE];
    if (st) {
        if (block_state(st, nullptr) != ST_RUNNING) return;
        if (prr && !fp_residual_test(prr, nb_rr, sc, st, status, lds)) return;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b32[i] = (float)r[i];
}

__global__ void k_f64_to_f32(size_t n, const double* __restrict__ in, float* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// Jacobi preconditioner z = D^-1 r with the r.z partials, optionally after the residual test
__global__ void __launch_bounds__(BLOCK)
k_fp_jacobi_prec(PcgArgs a, int h_next, int check) {
    __shared__ double lds[BLOCK / WAVE];
    if (block_state(a.st, nullptr) != ST_RUNNING) return;
    if (check && !fp_residual_test(a.prr, a.nb_rr, a.sc, a.st, a.status, lds)) return;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        const double zi = a.inv_diag[i] * a.r[i];
        a.z[i] = zi;
        acc += a.r[i] * zi;
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) a.prz[h_next][blockIdx.x] = acc;
}

// ======================================================================================================
// multicolour ILU(0) (src/cfa/linalg.cuh:361-568)
// ======================================================================================================
// Rows of one colour: for each elimination step of the row's schedule, l_ik = a_ik / u_kk and
// a_ij -= l_ik u_kj for the scheduled entries, in the schedule's order (colour by colour, CSR
// order inside a colour, then row k in column order: the order of linalg.cuh:366-392). Rows of
// one colour do not couple and rows k of earlier colours are final, so the update is in place.
__global__ void k_ilu_factor_rows(int m, int t0, const int* __restrict__ el_ptr, const int* __restrict__ el,
                                  const int* __restrict__ upd, double* __restrict__ lu) {
#pragma clang fp contract(off)   // separate multiply and subtract: the factors match a serial CPU run bit for bit
    const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= t0 + m) return;
    for (int e = el_ptr[t]; e < el_ptr[t + 1]; ++e) {
        const int q = el[4 * e], kd = el[4 * e + 1], u0 = el[4 * e + 2], u1 = el[4 * e + 3];
        const double lik = lu[q] / lu[kd];
        lu[q] = lik;
        for (int u = u0; u < u1; ++u) lu[upd[2 * u]] -= lik * lu[upd[2 * u + 1]];
    }
}

__global__ void k_ilu_gather(int nl, int nu, int n, const int* __restrict__ l_src, const int* __restrict__ u_src,
                             const int* __restrict__ d_src, const double* __restrict__ lu, double* __restrict__ lv,
                             double* __restrict__ uv, double* __restrict__ dinv, float* __restrict__ lv32,
                             float* __restrict__ uv32, float* __restrict__ dinv32) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nl) { const double v = lu[l_src[e]]; lv[e] = v; if (lv32) lv32[e] = (float)v; }
    if (e < nu) { const double v = lu[u_src[e]]; uv[e] = v; if (uv32) uv32[e] = (float)v; }
    if (e < n) { const double v = 1.0 / lu[d_src[e]]; dinv[e] = v; if (dinv32) dinv32[e] = (float)v; }
}

// The sliced-ELL values of the split rows after a factorization: slot k takes split entry src[k]
// (FP64 and, when kept, FP32 from the FP32 split values, so they equal the CSR path's), 0 for padding
__global__ void k_ilu_sell_values(int ls, int us, const int* __restrict__ lsrc, const int* __restrict__ usrc,
                                  const double* __restrict__ lv, const double* __restrict__ uv,
                                  const float* __restrict__ lv32, const float* __restrict__ uv32,
                                  double* __restrict__ slv, double* __restrict__ suv, float* __restrict__ slv32,
                                  float* __restrict__ suv32) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ls) {
        const int e = lsrc[k];
        slv[k] = e >= 0 ? lv[e] : 0.0;
        if (slv32) slv32[k] = e >= 0 ? lv32[e] : 0.0f;
    }
    if (k < us) {
        const int e = usrc[k];
        suv[k] = e >= 0 ? uv[e] : 0.0;
        if (suv32) suv32[k] = e >= 0 ? uv32[e] : 0.0f;
    }
}

// One colour's sweep on the sliced-ELL rows, one lane per row: forward (UPPER false: acc = r_i -
// sum l z, times dinv on the last colour) or backward (acc = z_i - sum u z, times dinv). All MAXL
// slots are loaded before the gathers; a padding slot gathers nothing and subtracts an exact +0,
// so each row's result is k_ilu_lower / k_ilu_upper's. mpos: the colour's positions (a multiple
// of 64) from global position p0.
template <int NRHS, int MAXL, bool UPPER, typename T>
__global__ void __launch_bounds__(BLOCK)
k_ilu_sweep_sell(int mpos, int n, int cs, int p0, const int* __restrict__ srow, const int* __restrict__ sp,
                 const int* __restrict__ sc, const T* __restrict__ sv, const T* __restrict__ dinv, int last,
                 const double* __restrict__ r, double* __restrict__ z, const int* __restrict__ st) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= mpos) return;
    int t, base, len;
    if (!sell_slot(sp, srow, p0 + p, st, t, base, len)) return;
    int c[MAXL];
    T v[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        const bool ok = k < len;
        c[k] = ok ? sc[base + SELL_SLICE * k] : 0;
        v[k] = ok ? sv[base + SELL_SLICE * k] : T(0);
    }
    if (t < 0) return;
    const int i = cs + t;
    double acc[NRHS];
#pragma unroll
    for (int q = 0; q < NRHS; ++q) acc[q] = UPPER ? z[(size_t)q * n + i] : r[(size_t)q * n + i];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        const double a = (double)v[k];
        const bool ok = k < len;
#pragma unroll
        for (int q = 0; q < NRHS; ++q) {
            const double zj = ok ? z[(size_t)q * n + c[k]] : 0.0;
            acc[q] -= a * zj;
        }
    }
    const double d = (double)dinv[t];
#pragma unroll
    for (int q = 0; q < NRHS; ++q) z[(size_t)q * n + i] = (UPPER || last) ? acc[q] * d : acc[q];
}

// The ILU(0) apply runs in colour order: r is permuted there (read in row order, written into one
// run per colour), the sweeps then stream every colour's rows, vector entries and factor entries,
// and z is permuted back (written in row order). In row order a colour's rows are spread over the
// whole vector, so each colour's sweep touched every cache line of r and z: measured at U7, 2.2x
// the algorithmic bytes crossed HBM.
// z back to row order (one colour: no backward sweep)
template <int NRHS>
__global__ void k_ilu_unpermute(int n, const int* __restrict__ pos, const double* __restrict__ zp,
                                double* __restrict__ z, const int* __restrict__ st) {
    if (halted(st)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = pos[i];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) z[(size_t)c * n + i] = zp[(size_t)c * n + q];
}

// The permutation into colour order fused with the forward sweep of colour 0. L holds the columns
// of earlier colours, so colour 0 has no L entries and its sweep is z = r (times dinv when it is
// also the last colour): its z entries are written here, and its r entries are not needed.
template <int NRHS, typename T>
__global__ void k_ilu_permute_first(int n, int m0, const int* __restrict__ pos, const double* __restrict__ r,
                                    const T* __restrict__ dinv, int last, double* __restrict__ rp,
                                    double* __restrict__ zp, const int* __restrict__ st) {
    if (halted(st)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = pos[i];
    if (q < m0) {
        const double d = (double)dinv[q];
#pragma unroll
        for (int c = 0; c < NRHS; ++c) {
            const double acc = r[(size_t)c * n + i];
            zp[(size_t)c * n + q] = last ? acc * d : acc;
        }
    } else {
#pragma unroll
        for (int c = 0; c < NRHS; ++c) rp[(size_t)c * n + q] = r[(size_t)c * n + i];
    }
}

// The backward sweep of colour 0 fused with the permutation back to row order, one thread per
// row i in row order: the rows of colour 0 (positions below m0) finish their sweep and write z
// directly, every other row copies its final value from z_perm.
template <int NRHS, typename T>
__global__ void __launch_bounds__(BLOCK)
k_ilu_last_unpermute(int n, int m0, const int* __restrict__ pos, const int* __restrict__ up,
                     const int* __restrict__ uc, const T* __restrict__ uv, const T* __restrict__ dinv,
                     const double* __restrict__ zp, double* __restrict__ z, const int* __restrict__ st) {
    if (halted(st)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int q = pos[i];
    if (q >= m0) {
#pragma unroll
        for (int c = 0; c < NRHS; ++c) z[(size_t)c * n + i] = zp[(size_t)c * n + q];
        return;
    }
    double acc[NRHS];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) acc[c] = zp[(size_t)c * n + q];
    row_entries<1>(uc, uv, up[q], up[q + 1], 0, [&](T ut, int j) {
        const double u = (double)ut;
#pragma unroll
        for (int c = 0; c < NRHS; ++c) acc[c] -= u * zp[(size_t)c * n + j];
    });
    const double d = (double)dinv[q];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) z[(size_t)c * n + i] = acc[c] * d;
}

// Forward sweep over the rows cs .. cs + m - 1 of one colour (colour order), all components per
// thread (the L row is read once). The row's entries are loaded ahead of their uses
// (row_entries) and subtracted in order, one lane per row: two lanes per row were 7-10 % faster
// at U6/U7 but sum in another order, which moves the reference's stored MMS error table in its
// last printed digit.
template <int NRHS, typename T>
__global__ void __launch_bounds__(BLOCK)
k_ilu_lower(int m, int n, int cs, const int* __restrict__ lp, const int* __restrict__ lc, const T* __restrict__ lv,
            const T* __restrict__ dinv, int last, const double* __restrict__ r, double* __restrict__ z,
            const int* __restrict__ st) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int e0, e1;
    if (!running_with_range(lp, t, m, st, e0, e1) || t >= m) return;
    const int i = cs + t;
    double acc[NRHS];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) acc[c] = r[(size_t)c * n + i];
    row_entries<1>(lc, lv, e0, e1, 0, [&](T lt, int j) {
        const double l = (double)lt;
#pragma unroll
        for (int c = 0; c < NRHS; ++c) acc[c] -= l * z[(size_t)c * n + j];
    });
    const double d = (double)dinv[t];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) z[(size_t)c * n + i] = last ? acc[c] * d : acc[c];
}

template <int NRHS, typename T>
__global__ void __launch_bounds__(BLOCK)
k_ilu_upper(int m, int n, int cs, const int* __restrict__ up, const int* __restrict__ uc, const T* __restrict__ uv,
            const T* __restrict__ dinv, double* __restrict__ z, const int* __restrict__ st) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int e0, e1;
    if (!running_with_range(up, t, m, st, e0, e1) || t >= m) return;
    const int i = cs + t;
    double acc[NRHS];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) acc[c] = z[(size_t)c * n + i];
    row_entries<1>(uc, uv, e0, e1, 0, [&](T ut, int j) {
        const double u = (double)ut;
#pragma unroll
        for (int c = 0; c < NRHS; ++c) acc[c] -= u * z[(size_t)c * n + j];
    });
    const double d = (double)dinv[t];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) z[(size_t)c * n + i] = acc[c] * d;
}

// The forward sweep with TPR lanes per row, for the FP32 factors of the mixed-precision path (no
// reference order to keep there): each lane sums its entries' products, the group adds them, and
// the row takes r - sum.
template <int NRHS, int TPR>
__global__ void __launch_bounds__(BLOCK)
k_ilu_lower_lanes(int m, int n, int cs, const int* __restrict__ lp, const int* __restrict__ lc,
                  const float* __restrict__ lv, const float* __restrict__ dinv, int last, const double* __restrict__ r,
                  double* __restrict__ z, const int* __restrict__ st) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = g / TPR, lane = g % TPR;
    int e0, e1;
    if (!running_with_range(lp, t, m, st, e0, e1)) return;
    const int i = cs + t;
    double sum[NRHS];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) sum[c] = 0.0;
    row_entries<TPR>(lc, lv, e0, e1, lane, [&](float lt, int j) {
        const double l = (double)lt;
#pragma unroll
        for (int c = 0; c < NRHS; ++c) sum[c] += l * z[(size_t)c * n + j];
    });
#pragma unroll
    for (int c = 0; c < NRHS; ++c) sum[c] = group_sum<TPR>(sum[c]);
    if (t >= m || lane != 0) return;
    const double d = (double)dinv[t];
#pragma unroll
    for (int c = 0; c < NRHS; ++c) {
        const double acc = r[(size_t)c * n + i] - sum[c];
        z[(size_t)c * n + i] = last ? acc * d : acc;
    }
}

// ======================================================================================================
// Two-component BiCGSTAB (src/cfa/linalg.cuh:642-779)
// ======================================================================================================
// Per-component scalars: slot(c) = base + c
enum {
    B_RN0 = 0, B_BD = 2, B_RHO = 4, B_RHO_OLD = 6, B_ALPHA = 8, B_OMEGA = 10, B_REL = 12,
    B_CB = 14, B_CO = 16, B_CA = 18, B_A1 = 20, B_A2 = 22, B_CW = 24, B_TOL = 26
};
// ints: [0] state (stopped when both components are inactive), [2 + c] active, [4 + c] iters,
// [6 + c] converged, [8 + c] half-step done, [10] current iteration, [11] iteration cap

// Two-component partial sums: part[b * 6 + 2 q + c]
// (lds: at least 2 nq * blockDim.x / 64 doubles)
template <int NQ>
__device__ __forceinline__ void write_partials2(double* part, double (&acc)[NQ][2], double* lds) {
    double v[2 * NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int c = 0; c < 2; ++c) v[2 * q + c] = acc[q][c];
    block_sum_k<2 * NQ>(v, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 2 * NQ; ++k) part[(size_t)blockIdx.x * 6 + k] = v[k];
}

// r = b - A x (both components), r0 = p = r, v = 0; partial r.r
template <int TPR>
__global__ void __launch_bounds__(BLOCK)
k_bicg_init(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ val,
            const double* __restrict__ b, const double* __restrict__ x, double* __restrict__ r,
            double* __restrict__ r0, double* __restrict__ p, double* __restrict__ v, double* __restrict__ part) {
    __shared__ double lds[2 * BLOCK / WAVE];
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int g0 = (blockIdx.x * blockDim.x + threadIdx.x) / TPR, lane = threadIdx.x % TPR;
    double acc[1][2] = {{0.0, 0.0}};
    for (int row = g0; row < n; row += groups) {
        double s0 = 0.0, s1 = 0.0;
        row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane, [&](double a, int j) {
            s0 += a * x[j];
            s1 += a * x[n + j];
        });
        s0 = group_sum<TPR>(s0);
        s1 = group_sum<TPR>(s1);
        if (lane == 0) {
            const double q0 = b[row] - 1.0 * s0, q1 = b[n + row] - 1.0 * s1;
            r[row] = q0;  r[n + row] = q1;
            r0[row] = q0; r0[n + row] = q1;
            p[row] = q0;  p[n + row] = q1;
            v[row] = 0.0; v[n + row] = 0.0;
            acc[0][0] += q0 * q0;
            acc[0][1] += q1 * q1;
        }
    }
    write_partials2<1>(part, acc, lds);
}

__device__ __forceinline__ void bicg_stop(int* st, int c, int it, int ok) {
    st[2 + c] = 0;
    st[4 + c] = it;
    st[6 + c] = ok;
}

__global__ void k_bicg_final_init(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                  int* __restrict__ st, int* __restrict__ status, int max_iter) {
    __shared__ double lds[2 * BLOCK / WAVE];
    double rr[2];
    sum_partials_k<2>({part, part + 1}, {nb, nb}, 6, rr, lds);
    if (threadIdx.x != 0) return;
    for (int c = 0; c < 2; ++c) {
        sc[B_RN0 + c] = sqrt(rr[c]);
        sc[B_RHO + c] = rr[c];
        sc[B_BD + c] = 1.0e-30 * sc[B_RN0 + c] * sc[B_RN0 + c] + 1.0e-300;
        sc[B_RHO_OLD + c] = 1.0;
        sc[B_ALPHA + c] = 1.0;
        sc[B_OMEGA + c] = 1.0;
        sc[B_REL + c] = 0.0;
        st[2 + c] = sc[B_RN0 + c] >= 1.0e-13 ? 1 : 0;
        st[4 + c] = 0;
        st[6 + c] = 1;
        st[8 + c] = 0;
    }
    st[10] = 0;
    st[11] = max_iter;
    st[0] = (st[2] || st[3]) && max_iter > 0 ? ST_RUNNING : ST_STOPPED;
    publish(status, 0, st[0]);
}

// Start of an iteration: breakdown check on rho, coefficients of the p update
__device__ void bicg_pre_logic(double* __restrict__ sc, int* __restrict__ st, int* __restrict__ status) {
    if (st[0] != ST_RUNNING) return;
    const int it = st[10];
    for (int c = 0; c < 2; ++c) {
        sc[B_CB + c] = 0.0;
        sc[B_CO + c] = 0.0;
        if (st[2 + c] && fabs(sc[B_RHO + c]) < sc[B_BD + c]) bicg_stop(st, c, it, 0);
        if (st[2 + c] && it > 0) {
            sc[B_CB + c] = (sc[B_RHO + c] / sc[B_RHO_OLD + c]) * (sc[B_ALPHA + c] / (sc[B_OMEGA + c] + 1.0e-30));
            sc[B_CO + c] = sc[B_OMEGA + c];
        }
    }
    if (!(st[2] || st[3])) { st[0] = ST_STOPPED; publish(status, it, ST_STOPPED); }
}

__global__ void k_bicg_pre(double* __restrict__ sc, int* __restrict__ st, int* __restrict__ status) {
    bicg_pre_logic(sc, st, status);
}

// p = r + cb (p - co v), from the second iteration on
__global__ void k_bicg_p(int n, const double* __restrict__ r, const double* __restrict__ v,
                         const double* __restrict__ sc, double* __restrict__ p, const int* __restrict__ st) {
    if (halted(st) || st[10] == 0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * n) return;
    const int c = i >= n;
    p[i] = r[i] + sc[B_CB + c] * (p[i] - sc[B_CO + c] * v[i]);
}

// y = A x for both components; with dots: partial y.w0 (mode 1) or y.y, y.w0 (mode 2)
template <int TPR>
__global__ void __launch_bounds__(BLOCK)
k_spmv2_dots(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ val,
             const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ w0, int mode,
             double* __restrict__ part, const int* __restrict__ st) {
    __shared__ double lds[4 * BLOCK / WAVE];
    if (halted(st)) return;
    const int groups = gridDim.x * (blockDim.x / TPR);
    const int g0 = (blockIdx.x * blockDim.x + threadIdx.x) / TPR, lane = threadIdx.x % TPR;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int row = g0; row < n; row += groups) {
        double s0 = 0.0, s1 = 0.0;
        row_entries<TPR>(ci, val, rp[row], rp[row + 1], lane, [&](double a, int j) {
            s0 += a * x[j];
            s1 += a * x[n + j];
        });
        s0 = group_sum<TPR>(s0);
        s1 = group_sum<TPR>(s1);
        if (lane == 0) {
            y[row] = s0;
            y[n + row] = s1;
            if (mode == 1) {   // r0 . v
                acc[0][0] += w0[row] * s0;
                acc[0][1] += w0[n + row] * s1;
            } else {           // t . t, t . s
                acc[0][0] += s0 * s0;
                acc[0][1] += s1 * s1;
                acc[1][0] += s0 * w0[row];
                acc[1][1] += s1 * w0[n + row];
            }
        }
    }
    if (mode == 1) write_partials2<1>(part, reinterpret_cast<double(&)[1][2]>(acc), lds);
    else write_partials2<2>(part, acc, lds);
}

// The scalar steps between the vector kernels are folded into the kernels that consume them:
// every block sums the producer's partials in the same fixed order and takes the same decision;
// block 0 alone writes the state. A component stopped here is treated by a late block exactly
// as an already inactive one (zero coefficient), so a block that reads the state before or after
// block 0's write computes the same values.

// alpha = rho / (r0 . v) with the breakdown guard (from the r0.v partials), then s = r - alpha v
// and the s.s partials
__global__ void __launch_bounds__(BLOCK)
k_bicg_s(int n, const double* __restrict__ r, const double* __restrict__ v, double* __restrict__ sc,
         double* __restrict__ s, double* __restrict__ part, int nb_a, const double* __restrict__ part_a,
         int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[2 * BLOCK / WAVE];
    __shared__ int s_state, s_act[2];
    if (threadIdx.x == 0) { s_state = st[0]; s_act[0] = st[2]; s_act[1] = st[3]; }
    __syncthreads();
    if (s_state != ST_RUNNING) return;
    double ca[2];
    bool act[2];
    double o2[2];
    sum_partials_k<2>({part_a, part_a + 1}, {nb_a, nb_a}, 6, o2, lds);
    for (int c = 0; c < 2; ++c) {
        const double o = o2[c];
        act[c] = s_act[c] && !(fabs(o) < sc[B_BD + c]);
        ca[c] = act[c] ? sc[B_RHO + c] / o : 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int it = st[10];
        for (int c = 0; c < 2; ++c) {
            sc[B_CA + c] = ca[c];
            if (s_act[c] && !act[c]) bicg_stop(st, c, it, 0);
            if (act[c]) sc[B_ALPHA + c] = ca[c];
        }
        if (!(act[0] || act[1])) { st[0] = ST_STOPPED; publish(status, it, ST_STOPPED); }
    }
    if (!(act[0] || act[1])) return;
    double acc[1][2] = {{0.0, 0.0}};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += gridDim.x * blockDim.x) {
        const int c = i >= n;
        const double si = r[i] - ca[c] * v[i];
        s[i] = si;
        acc[0][c] += si * si;
    }
    write_partials2<1>(part, acc, lds);
}

// omega = (t . s) / (t . t) unless the half step already converged (from the s.s and t.t, t.s
// partials), then x += alpha phat + omega shat, r = s - omega t and the r.r, r0.r partials
__global__ void __launch_bounds__(BLOCK)
k_bicg_xr(int n, const double* __restrict__ ph, const double* __restrict__ sh, const double* __restrict__ s,
          const double* __restrict__ t, const double* __restrict__ r0, double* __restrict__ sc, double* __restrict__ x,
          double* __restrict__ r, double* __restrict__ part, int nb_s, const double* __restrict__ part_s, int nb_t,
          const double* __restrict__ part_t, int* __restrict__ st) {
    __shared__ double lds[6 * BLOCK / WAVE];
    __shared__ int s_state, s_act[2];
    if (threadIdx.x == 0) { s_state = st[0]; s_act[0] = st[2]; s_act[1] = st[3]; }
    __syncthreads();
    if (s_state != ST_RUNNING) return;
    double a1[2], a2[2], om[2];
    int half[2];
    double sums[6];   // s.s, t.t, t.s of both components
    sum_partials_k<6>({part_s, part_s + 1, part_t, part_t + 1, part_t + 2, part_t + 3},
                      {nb_s, nb_s, nb_t, nb_t, nb_t, nb_t}, 6, sums, lds);
    for (int c = 0; c < 2; ++c) {
        const double ss = sums[c];
        const double tt = sums[2 + c];
        const double ts = sums[4 + c];
        a1[c] = a2[c] = om[c] = 0.0;
        half[c] = 0;
        if (!s_act[c]) continue;
        a1[c] = sc[B_ALPHA + c];
        const double s_norm = sqrt(ss);
        if (s_norm < sc[B_TOL] * sc[B_RN0 + c] || s_norm < 1.0e-11) {
            half[c] = 1;
            continue;
        }
        om[c] = (tt > sc[B_BD + c]) ? (ts / tt) : 0.0;
        a2[c] = om[c];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int c = 0; c < 2; ++c) {
            sc[B_A1 + c] = a1[c];
            sc[B_A2 + c] = a2[c];
            sc[B_CW + c] = a2[c];
            st[8 + c] = half[c];
            if (s_act[c] && !half[c]) sc[B_OMEGA + c] = om[c];
        }
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += gridDim.x * blockDim.x) {
        const int c = i >= n;
        x[i] += a1[c] * ph[i] + a2[c] * sh[i];
        const double ri = s[i] - a2[c] * t[i];
        r[i] = ri;
        acc[0][c] += ri * ri;
        acc[1][c] += r0[i] * ri;
    }
    write_partials2<2>(part, acc, lds);
}

// end of an iteration: half-step stops, residual test, rho update, iteration cap; then the start
// of the next one (breakdown check on rho, p-update coefficients)
__global__ void k_bicg_final_post(int nb, const double* __restrict__ part, double* __restrict__ sc,
                                  int* __restrict__ st, int* __restrict__ status) {
    __shared__ double lds[4 * BLOCK / WAVE];
    if (st[0] != ST_RUNNING) return;
    double sums[4];
    sum_partials_k<4>({part, part + 1, part + 2, part + 3}, {nb, nb, nb, nb}, 6, sums, lds);
    const double rr[2] = {sums[0], sums[1]}, r0r[2] = {sums[2], sums[3]};
    if (threadIdx.x != 0) return;
    const int it = st[10], max_iter = st[11];
    for (int c = 0; c < 2; ++c)
        if (st[8 + c]) bicg_stop(st, c, it + 1, 1);
    for (int c = 0; c < 2; ++c) {
        if (!st[2 + c]) continue;
        const double r_norm = sqrt(rr[c]);
        sc[B_REL + c] = r_norm / sc[B_RN0 + c];
        if (r_norm < sc[B_TOL] * sc[B_RN0 + c] || r_norm < 1.0e-11) { bicg_stop(st, c, it + 1, 1); continue; }
        sc[B_RHO_OLD + c] = sc[B_RHO + c];
        sc[B_RHO + c] = r0r[c];
        if (it + 1 == max_iter) bicg_stop(st, c, max_iter, 0);
    }
    st[10] = it + 1;
    const int state = (st[2] || st[3]) ? ST_RUNNING : ST_STOPPED;
    st[0] = state;
    publish(status, it + 1, state);
    bicg_pre_logic(sc, st, status);   // start of the next iteration
}

// ======================================================================================================
// Host side
// ======================================================================================================
hipStream_t S(cfa_stream_t s) { return as_stream(s); }

#define CFA_TRY(expr) do { const hipError_t e_ = (expr); if (e_ != hipSuccess) return (int)e_; } while (0)

// Calls f with std::integral_constant<int, TPR> for the lanes-per-row value tpr (1..64)
template <typename F>
void tpr_dispatch(int tpr, F f) {
    switch (tpr) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        default: f(std::integral_constant<int, 64>{}); break;
    }
}

template <int OP, typename T = double>
void csr_op_launch(int tpr, int rows, const int* rp, const int* ci, const T* val, const T* x, const T* b,
                   const T* dinv, double omega, T* y, const int* st, hipStream_t s) {
    const int blocks = blocks_for((long long)rows * tpr);
    tpr_dispatch(tpr, [&](auto c) {
        constexpr int TPR = decltype(c)::value;
        hipLaunchKernelGGL((k_csr_op<TPR, OP, T>), dim3(blocks), dim3(BLOCK), 0, s, rows, rp, ci, val, x, b, dinv,
                           (T)omega, y, st);
    });
}

template <typename T = double>
void jacobi_from_zero_launch(int tpr, int rows, const int* rp, const int* ci, const T* val, const T* b, const T* dinv,
                             double omega, T* y, const int* st, hipStream_t s) {
    const int blocks = blocks_for((long long)rows * tpr);
    tpr_dispatch(tpr, [&](auto c) {
        constexpr int TPR = decltype(c)::value;
        hipLaunchKernelGGL((k_jacobi_from_zero<TPR, T>), dim3(blocks), dim3(BLOCK), 0, s, rows, rp, ci, val, b, dinv,
                           (T)omega, y, st);
    });
}

// Hooks of the fused PCG into the preconditioner: the iteration's residual test runs in the first
// level-0 kernel, and the last level-0 sweep emits the r.z partials.
struct PrecHooks {
    const double* prr = nullptr;   // r.r partials of the iteration (residual test when set)
    int nb_rr = 0;
    double* sc = nullptr;
    int* st_w = nullptr;
    int* status = nullptr;
    double* dot_part = nullptr;    // r.z partials out (when set)
    int dot_blocks = 0;            // set: number of r.z partials written
};

// last level-0 sweep into the FP64 output, with the r.z partials when part is set; returns the
// number of partials
template <typename T = double>
int jacobi_dot_launch(int tpr, int rows, const int* rp, const int* ci, const T* val, const T* x, const T* b,
                      const T* dinv, double omega, double* y, const double* rdot, double* part, const int* st,
                      hipStream_t s) {
    int nb = 0;
    tpr_dispatch(tpr, [&](auto c) {
        constexpr int TPR = decltype(c)::value;
        nb = std::min(blocks_for((long long)rows * TPR), RED_BLOCKS);
        hipLaunchKernelGGL((k_jacobi_dot<TPR, T>), dim3(nb), dim3(BLOCK), 0, s, rows, rp, ci, val, x, b, dinv,
                           (T)omega, y, rdot, part, st);
    });
    return nb;
}

void jacobi_from_zero_check_launch(int tpr, int rows, const int* rp, const int* ci, const double* val,
                                   const double* b, const double* dinv, double omega, double* y,
                                   const PrecHooks& hk, hipStream_t s) {
    const int blocks = blocks_for((long long)rows * tpr);
    tpr_dispatch(tpr, [&](auto c) {
        constexpr int TPR = decltype(c)::value;
        hipLaunchKernelGGL(k_jacobi_from_zero_check<TPR>, dim3(blocks), dim3(BLOCK), 0, s, rows, rp, ci, val, b, dinv,
                           omega, y, hk.prr, hk.nb_rr, hk.sc, hk.st_w, hk.status);
    });
}

// MAXL (slots loaded per lane) of the sliced-ELL kernels for the level's longest row (at most
// CFA_SELL_MAX_LEN), G, the CSR lanes per row whose summation order they keep (1, 2 or 4), and NT,
// non-temporal loads when a sweep's bytes (values of vb bytes) exceed the Infinity Cache
template <typename F>
void sell_dispatch(const cfa_amg_level_t& L, int vb, F f) {
    const bool nt = (4.0 + vb) * L.a_nnz + (4.0 + 4.0 * vb) * L.rows > SELL_STREAM_BYTES;
    auto with_g = [&](auto m) {
        auto with_nt = [&](auto g) {
            if (nt) f(m, g, std::true_type{});
            else f(m, g, std::false_type{});
        };
        if (L.tpr_a == 1) with_nt(std::integral_constant<int, 1>{});
        else if (L.tpr_a == 2) with_nt(std::integral_constant<int, 2>{});
        else with_nt(std::integral_constant<int, 4>{});
    };
    const int maxl = L.s_len;
    if (maxl <= 8) with_g(std::integral_constant<int, 8>{});
    else if (maxl <= 12) with_g(std::integral_constant<int, 12>{});
    else with_g(std::integral_constant<int, 16>{});
}

// A level's sliced-ELL copy as the kernels take it
bool sell_usable(const cfa_amg_level_t& L) {
    return L.s_row && L.s_ci && L.s_v && L.s_len >= 0 && L.s_len <= CFA_SELL_MAX_LEN &&
           (L.tpr_a == 1 || L.tpr_a == 2 || L.tpr_a == 4);
}

// The level operator's sweeps and residual: on its sliced-ELL copy when it has one, else on its
// CSR rows
template <int OP, typename T>
void level_op_launch(const cfa_amg_level_t& L, const T* x, const T* b, const T* dinv, T* y, const int* st,
                     hipStream_t s) {
    static_assert(OP == OP_JACOBI || OP == OP_RESID, "the sliced-ELL copy serves the sweeps and the residual");
    if (!L.s_sp) {
        csr_op_launch<OP, T>(L.tpr_a, L.rows, L.a_rp, L.a_ci, Lv<T>::a(L), x, b, dinv, OP == OP_JACOBI ? L.omega : 0.0,
                             y, st, s);
        return;
    }
    const int pos = sell_positions(L.rows);
    sell_dispatch(L, sizeof(T), [&](auto m, auto g, auto nt) {
        constexpr int M = decltype(m)::value, G = decltype(g)::value;
        constexpr bool NT = decltype(nt)::value;
        hipLaunchKernelGGL((k_sell_op<M, G, NT, OP, T>), dim3(blocks_for(pos)), dim3(BLOCK), 0, s, pos, L.s_sp, L.s_row,
                           L.s_ci, Lv<T>::sv(L), x, b, dinv, (T)L.omega, y, st);
    });
}

template <typename T>
void level_from_zero_launch(const cfa_amg_level_t& L, const T* b, const T* dinv, T* y, const int* st, hipStream_t s) {
    if (!L.s_sp) {
        jacobi_from_zero_launch<T>(L.tpr_a, L.rows, L.a_rp, L.a_ci, Lv<T>::a(L), b, dinv, L.omega, y, st, s);
        return;
    }
    const int pos = sell_positions(L.rows);
    sell_dispatch(L, sizeof(T), [&](auto m, auto g, auto nt) {
        constexpr int M = decltype(m)::value, G = decltype(g)::value;
        constexpr bool NT = decltype(nt)::value;
        hipLaunchKernelGGL((k_sell_from_zero<M, G, NT, T>), dim3(blocks_for(pos)), dim3(BLOCK), 0, s, pos, L.s_sp,
                           L.s_row, L.s_ci, Lv<T>::sv(L), b, dinv, (T)L.omega, y, st);
    });
}

void level_from_zero_check_launch(const cfa_amg_level_t& L, const double* b, double* y, const PrecHooks& hk,
                                  hipStream_t s) {
    if (!L.s_sp) {
        jacobi_from_zero_check_launch(L.tpr_a, L.rows, L.a_rp, L.a_ci, L.a_v, b, L.dinv, L.omega, y, hk, s);
        return;
    }
    const int pos = sell_positions(L.rows);
    sell_dispatch(L, sizeof(double), [&](auto m, auto g, auto nt) {
        constexpr int M = decltype(m)::value, G = decltype(g)::value;
        constexpr bool NT = decltype(nt)::value;
        hipLaunchKernelGGL((k_sell_from_zero_check<M, G, NT>), dim3(blocks_for(pos)), dim3(BLOCK), 0, s, pos, L.s_sp,
                           L.s_row, L.s_ci, L.s_v, b, L.dinv, L.omega, y, hk.prr, hk.nb_rr, hk.sc, hk.st_w, hk.status);
    });
}

// One V-cycle (src/cfa/linalg.cuh:1184-1206) for level l with right-hand side b, in FP64 or FP32
// (T). Each level's x / ax buffers form a ping-pong pair; the function returns the buffer holding
// the level's iterate. On level 0 the last smoothing step writes straight into the FP64 `out`
// when given (rdot: the FP64 residual for the r.z partials).
template <typename T>
T* vcycle(const cfa_amg_t* amg, int l, const T* b, double* out, const double* rdot, const int* st, hipStream_t s,
          PrecHooks* hk = nullptr) {
    const cfa_amg_level_t& L = amg->levels[l];
    const int n = L.rows;
    const int ss = amg->smooth_steps;
    T* xa = Lv<T>::x(L);
    T* xb = Lv<T>::ax(L);
    const T* dv = Lv<T>::dinv(L);
    // the residual test of a fused PCG iteration (the FP32 cycle takes it in its entry kernel)
    const bool check_here = std::is_same<T, double>::value && hk && hk->prr && l == 0;
    if (ss >= 2) {
        bool done = false;
        if constexpr (std::is_same<T, double>::value) {
            if (check_here) {
                level_from_zero_check_launch(L, b, xa, *hk, s);
                done = true;
            }
        }
        if (!done) level_from_zero_launch<T>(L, b, dv, xa, st, s);
        for (int k = 2; k < ss; ++k) {
            level_op_launch<OP_JACOBI, T>(L, xa, b, dv, xb, st, s);
            std::swap(xa, xb);
        }
    } else {   // x = omega dinv b is always taken, also for fewer than one smoothing step
        if (check_here)
            hipLaunchKernelGGL(k_fp_check, dim3(1), dim3(BLOCK), 0, s, hk->prr, hk->nb_rr, hk->sc, hk->st_w, hk->status);
        hipLaunchKernelGGL(k_jacobi_init<T>, dim3(blocks_for(n)), dim3(BLOCK), 0, s, n, (T)L.omega, dv, b, xa, st);
    }
    level_op_launch<OP_RESID, T>(L, xa, b, nullptr, Lv<T>::res(L), st, s);
    const bool deeper = l + 1 < amg->n_levels;
    T* cb = deeper ? Lv<T>::b(amg->levels[l + 1]) : Lv<T>::cb(amg);
    csr_op_launch<OP_SPMV, T>(L.tpr_r, L.cols, L.r_rp, L.r_ci, Lv<T>::r(L), Lv<T>::res(L), nullptr, nullptr, 0.0, cb,
                              st, s);
    const T* cx;
    if (deeper) {
        cx = vcycle<T>(amg, l + 1, Lv<T>::b(amg->levels[l + 1]), nullptr, nullptr, st, s);
    } else {
        hipLaunchKernelGGL(k_dense_matvec<T>, dim3(blocks_for((long long)amg->coarse_n * WAVE)), dim3(BLOCK), 0, s,
                           amg->coarse_n, amg->coarse_inv, Lv<T>::cb(amg), Lv<T>::cx(amg), st);
        cx = Lv<T>::cx(amg);
    }
    csr_op_launch<OP_ADD, T>(L.tpr_p, n, L.p_rp, L.p_ci, Lv<T>::p(L), cx, nullptr, nullptr, 0.0, xa, st, s);
    for (int k = 0; k < ss; ++k) {
        if (out && k == ss - 1) {
            // on the CSR rows: its r.z block partials keep their grouping by rows
            const int nb = jacobi_dot_launch<T>(L.tpr_a, n, L.a_rp, L.a_ci, Lv<T>::a(L), xa, b, dv, L.omega, out,
                                                rdot, hk ? hk->dot_part : nullptr, st, s);
            if (hk && hk->dot_part) hk->dot_blocks = nb;
            return nullptr;   // the result is in out
        }
        level_op_launch<OP_JACOBI, T>(L, xa, b, dv, xb, st, s);
        std::swap(xa, xb);
    }
    return xa;
}

// A hierarchy descriptor the V-cycle can run: sizes, and the FP32 arrays when use_f32 is set
bool amg_valid(const cfa_amg_t* amg) {
    if (amg->n_levels < 0 || (amg->n_levels > 0 && !amg->levels) || amg->coarse_n < 0 || amg->smooth_steps < 0)
        return false;
    if (amg->use_f32) {
        if (amg->n_levels > 0 && (!amg->cb32 || !amg->cx32)) return false;
        for (int l = 0; l < amg->n_levels; ++l) {
            const cfa_amg_level_t& L = amg->levels[l];
            if (!L.a_v32 || !L.p_v32 || !L.r_v32 || !L.dinv32 || !L.x32 || !L.b32 || !L.r32 || !L.ax32) return false;
            if (L.s_sp && !L.s_v32) return false;
        }
    }
    for (int l = 0; l < amg->n_levels; ++l) {
        const cfa_amg_level_t& L = amg->levels[l];
        if (L.s_sp && !sell_usable(L)) return false;
    }
    return true;
}

// z = V-cycle(r); with hooks, the residual test first and the r.z partials last. The FP32 cycle
// (amg->use_f32) reads r through b32 and writes z in FP64.
void amg_apply_enqueue(const cfa_amg_t* amg, const double* r, double* z, const int* st, hipStream_t s,
                       PrecHooks* hk = nullptr) {
    const int n = amg->n_levels ? amg->levels[0].rows : amg->coarse_n;
    const bool f32 = amg->use_f32 && amg->n_levels > 0 && amg->smooth_steps >= 1;
    bool wrote_z = false;
    if (amg->n_levels == 0) {
        if (hk && hk->prr)
            hipLaunchKernelGGL(k_fp_check, dim3(1), dim3(BLOCK), 0, s, hk->prr, hk->nb_rr, hk->sc, hk->st_w, hk->status);
        hipLaunchKernelGGL(k_dense_matvec<double>, dim3(blocks_for((long long)amg->coarse_n * WAVE)), dim3(BLOCK), 0, s,
                           amg->coarse_n, amg->coarse_inv, r, z, st);
    } else if (f32) {
        float* b32 = amg->levels[0].b32;
        hipLaunchKernelGGL(k_to_f32_check, dim3(red_blocks(n)), dim3(BLOCK), 0, s, n, r, b32,
                           hk ? hk->prr : nullptr, hk ? hk->nb_rr : 0, hk ? hk->sc : nullptr,
                           hk && hk->prr ? hk->st_w : const_cast<int*>(st), hk ? hk->status : nullptr);
        (void)vcycle<float>(amg, 0, b32, z, r, st, s, hk);
        wrote_z = true;
    } else {
        double* res = vcycle<double>(amg, 0, r, z, r, st, s, hk);
        wrote_z = res == nullptr;
        if (!wrote_z) hipLaunchKernelGGL(k_copy, dim3(blocks_for(n)), dim3(BLOCK), 0, s, n, res, z, st);
    }
    if (hk && hk->dot_part && !wrote_z) {
        hk->dot_blocks = red_blocks(n);
        hipLaunchKernelGGL(k_dot_partial, dim3(hk->dot_blocks), dim3(BLOCK), 0, s, n, r, z, hk->dot_part, st);
    }
}

template <int NRHS, typename T>
void ilu_apply_impl(const cfa_ilu_t* ilu, const T* lv, const T* uv, const T* dinv, const double* r, double* z,
                    const int* st, hipStream_t s) {
    const int nc = ilu->n_colours, n = ilu->n;
    if (n == 0 || nc == 0) return;
    double* rp = ilu->r_perm;
    double* zp = ilu->z_perm;
    const int m0 = ilu->colour_start[1] - ilu->colour_start[0];
    // colour 0 forward with the permutation, colours 1.. forward, colours nc-2..1 backward, colour 0
    // backward with the permutation back (the last colour has no U entries)
    hipLaunchKernelGGL((k_ilu_permute_first<NRHS, T>), dim3(blocks_for(n)), dim3(BLOCK), 0, s, n, m0, ilu->pos, r,
                       dinv, nc == 1 ? 1 : 0, rp, zp, st);
    // forward sweeps with the FP32 factors (mixed precision): two lanes per row; FP64: one lane, the
    // reference's order (the backward sweeps take one lane either way: two were no faster)
    constexpr bool lanes = std::is_same<T, float>::value;
    // the one-lane sweeps on the sliced-ELL rows when the factor has them
    const T* slv = std::is_same<T, float>::value ? (const T*)(const void*)ilu->s_lv32 : (const T*)(const void*)ilu->s_lv;
    const T* suv = std::is_same<T, float>::value ? (const T*)(const void*)ilu->s_uv32 : (const T*)(const void*)ilu->s_uv;
    const bool sell_l = ilu->s_lrow && slv && ilu->s_pos0, sell_u = ilu->s_urow && suv && ilu->s_pos0;
    auto sell_sweep = [&](auto upper, int c, int len, const int* srow, const int* sp, const int* sc, const T* sv) {
        const int cs = ilu->colour_start[c], mpos = (ilu->colour_start[c + 1] - cs + 63) / 64 * 64;
        auto go = [&](auto mx) {
            constexpr int M = decltype(mx)::value;
            constexpr bool U = decltype(upper)::value;
            hipLaunchKernelGGL((k_ilu_sweep_sell<NRHS, M, U, T>), dim3(blocks_for(mpos)), dim3(BLOCK), 0, s, mpos, n,
                               cs, ilu->s_pos0[c], srow, sp, sc, sv, dinv + cs, (!U && c == nc - 1) ? 1 : 0, rp, zp,
                               st);
        };
        if (len <= 8) go(std::integral_constant<int, 8>{});
        else if (len <= 12) go(std::integral_constant<int, 12>{});
        else go(std::integral_constant<int, 16>{});
    };
    for (int c = 1; c < nc; ++c) {
        const int cs = ilu->colour_start[c], m = ilu->colour_start[c + 1] - cs;
        if (m == 0) continue;
        if (!lanes && sell_l)
            sell_sweep(std::false_type{}, c, ilu->s_llen, ilu->s_lrow, ilu->s_lsp, ilu->s_lc, slv);
        else if constexpr (lanes)
            hipLaunchKernelGGL((k_ilu_lower_lanes<NRHS, 2>), dim3(blocks_for(2LL * m)), dim3(BLOCK), 0, s, m, n, cs,
                               ilu->lp + cs, ilu->lc, lv, dinv + cs, c == nc - 1 ? 1 : 0, rp, zp, st);
        else
            hipLaunchKernelGGL((k_ilu_lower<NRHS, T>), dim3(blocks_for(m)), dim3(BLOCK), 0, s, m, n, cs, ilu->lp + cs,
                               ilu->lc, lv, dinv + cs, c == nc - 1 ? 1 : 0, rp, zp, st);
    }
    for (int c = nc - 2; c >= 1; --c) {
        const int cs = ilu->colour_start[c], m = ilu->colour_start[c + 1] - cs;
        if (m == 0) continue;
        if (sell_u) {
            sell_sweep(std::true_type{}, c, ilu->s_ulen, ilu->s_urow, ilu->s_usp, ilu->s_uc, suv);
            continue;
        }
        hipLaunchKernelGGL((k_ilu_upper<NRHS, T>), dim3(blocks_for(m)), dim3(BLOCK), 0, s, m, n, cs, ilu->up + cs,
                           ilu->uc, uv, dinv + cs, zp, st);
    }
    if (nc >= 2)
        hipLaunchKernelGGL((k_ilu_last_unpermute<NRHS, T>), dim3(blocks_for(n)), dim3(BLOCK), 0, s, n, m0, ilu->pos,
                           ilu->up, ilu->uc, uv, dinv, zp, z, st);
    else
        hipLaunchKernelGGL((k_ilu_unpermute<NRHS>), dim3(blocks_for(n)), dim3(BLOCK), 0, s, n, ilu->pos, zp, z, st);
}

void ilu_apply_any(const cfa_ilu_t* ilu, int use_f32, const double* r, double* z, int nrhs, const int* st,
                   hipStream_t s) {
    if (use_f32 && ilu->lv32) {
        if (nrhs == 2) ilu_apply_impl<2, float>(ilu, ilu->lv32, ilu->uv32, ilu->dinv32, r, z, st, s);
        else ilu_apply_impl<1, float>(ilu, ilu->lv32, ilu->uv32, ilu->dinv32, r, z, st, s);
    } else {
        if (nrhs == 2) ilu_apply_impl<2, double>(ilu, ilu->lv, ilu->uv, ilu->dinv, r, z, st, s);
        else ilu_apply_impl<1, double>(ilu, ilu->lv, ilu->uv, ilu->dinv, r, z, st, s);
    }
}

// Host wait on the progress word: returns when progress >= target, the solve stopped, or the
// stream is idle.
void wait_progress(const cfa_krylov_ws_t* ws, int target, hipStream_t s) {
    volatile int* st = ws->status_host;
    for (int spin = 0;; ++spin) {
        if (st[0] >= target || st[1] != ST_RUNNING) return;
        if ((spin & 63) == 63) {
            if (hipStreamQuery(s) == hipSuccess) return;
            std::this_thread::yield();
        }
    }
}

int* state_words(cfa_krylov_ws_t* ws) { return reinterpret_cast<int*>(ws->scal + 40); }

// Captured iteration graphs, keyed by everything the captured launches depend on
struct IterGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    unsigned long long used = 0;
};
struct WsInternal {
    hipStream_t cap = nullptr;
    IterGraph pcg[GRAPH_CACHE], bicg[GRAPH_CACHE];
    unsigned long long clock = 0;
    bool capture_broken = false;
};

WsInternal* internal_of(cfa_krylov_ws_t* ws) {
    if (!ws->internal) ws->internal = new WsInternal();
    return static_cast<WsInternal*>(ws->internal);
}

struct KeyBuilder {
    std::vector<unsigned char> k;
    template <typename T>
    void add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        k.insert(k.end(), p, p + sizeof(T));
    }
    void add_bytes(const void* p, size_t n) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        k.insert(k.end(), q, q + n);
    }
};

// The cached graph for `key` (least recently used slot replaced), captured from `enqueue`;
// nullptr when capture is unavailable (the caller then launches directly).
template <typename F>
hipGraphExec_t iteration_graph(cfa_krylov_ws_t* ws, IterGraph* cache, const std::vector<unsigned char>& key,
                               F enqueue) {
    WsInternal* in = internal_of(ws);
    if (!ws->use_graphs || in->capture_broken) return nullptr;
    IterGraph* slot = &cache[0];
    for (int k = 0; k < GRAPH_CACHE; ++k) {
        if (cache[k].exec && cache[k].key == key) {
            cache[k].used = ++in->clock;
            return cache[k].exec;
        }
        if (cache[k].used < slot->used) slot = &cache[k];
    }
    if (slot->exec) { (void)hipGraphExecDestroy(slot->exec); slot->exec = nullptr; }
    if (!in->cap && hipStreamCreateWithFlags(&in->cap, hipStreamNonBlocking) != hipSuccess) {
        in->cap = nullptr;
        in->capture_broken = true;
        (void)hipGetLastError();
        return nullptr;
    }
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(in->cap, hipStreamCaptureModeRelaxed) == hipSuccess;
    if (ok) {
        enqueue(in->cap);
        ok = hipStreamEndCapture(in->cap, &graph) == hipSuccess && graph;
    }
    if (ok) ok = hipGraphInstantiate(&slot->exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!ok) {
        slot->exec = nullptr;
        in->capture_broken = true;
        return nullptr;
    }
    slot->key = key;
    slot->used = ++in->clock;
    return slot->exec;
}

// The iteration loop shared by both solvers: enqueue iterations (graph replay or direct
// launches), staying LOOKAHEAD iterations ahead of the progress the device has published.
template <typename F>
int run_iterations(cfa_krylov_ws_t* ws, hipGraphExec_t graph, int max_iter, F enqueue, hipStream_t s) {
    for (int it = 0; it < max_iter; ++it) {
        if (it >= LOOKAHEAD) {
            wait_progress(ws, it - LOOKAHEAD + 1, s);
            if (ws->status_host[1] != ST_RUNNING) break;
        }
        if (graph) CFA_TRY(hipGraphLaunch(graph, s));
        else enqueue(s);
    }
    return launch_status();
}

}  // namespace

extern "C" {

int cfa_krylov_ws_release(cfa_krylov_ws_t* ws) {
    if (!ws || !ws->internal) return CFA_OK;
    WsInternal* in = static_cast<WsInternal*>(ws->internal);
    for (int k = 0; k < GRAPH_CACHE; ++k) {
        if (in->pcg[k].exec) (void)hipGraphExecDestroy(in->pcg[k].exec);
        if (in->bicg[k].exec) (void)hipGraphExecDestroy(in->bicg[k].exec);
    }
    if (in->cap) (void)hipStreamDestroy(in->cap);
    delete in;
    ws->internal = nullptr;
    return CFA_OK;
}

int cfa_spmv(int n, const int* rp, const int* ci, const double* val, const double* x, double* y, int nrhs,
             cfa_stream_t stream) {
    if (n < 0 || !rp || !ci || !val || !x || !y || nrhs < 1 || nrhs > 2) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    // the mesh matrices (5-9 entries per row): two lanes per row from 20k rows (measured fastest
    // for the level-0 sweeps of the same pattern), a group of eight on small, latency-bound meshes
    auto launch = [&](auto c) {
        constexpr int TPR = decltype(c)::value;
        const int blocks = blocks_for((long long)n * TPR);
        if (nrhs == 1)
            hipLaunchKernelGGL((k_spmv_multi<TPR, 1>), dim3(blocks), dim3(BLOCK), 0, S(stream), n, rp, ci, val, x, y);
        else
            hipLaunchKernelGGL((k_spmv_multi<TPR, 2>), dim3(blocks), dim3(BLOCK), 0, S(stream), n, rp, ci, val, x, y);
    };
    if (n >= 20000) launch(std::integral_constant<int, 2>{});
    else launch(std::integral_constant<int, 8>{});
    return launch_status();
}

int cfa_dot(int n, const double* x, const double* y, double* result, double* ws, cfa_stream_t stream) {
    if (n < 0 || !x || !y || !result || !ws) return CFA_E_ARG;
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, S(stream), n, x, y, ws, nullptr);
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(BLOCK), 0, S(stream), nb, ws, ws + RED_BLOCKS, 0);
    const int rc = launch_status();
    if (rc) return rc;
    CFA_TRY(hipMemcpyAsync(result, ws + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
    CFA_TRY(hipStreamSynchronize(S(stream)));
    return CFA_OK;
}

int cfa_axpy(int n, double a, const double* x, double* y, cfa_stream_t stream) {
    if (n < 0 || !x || !y) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_axpy, dim3(blocks_for(n)), dim3(BLOCK), 0, S(stream), n, a, x, y);
    return launch_status();
}

int cfa_xpay(int n, const double* x, double b, double* y, cfa_stream_t stream) {
    if (n < 0 || !x || !y) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_xpay, dim3(blocks_for(n)), dim3(BLOCK), 0, S(stream), n, x, b, y);
    return launch_status();
}

int cfa_vec_mul(int n, const double* a, const double* b, double* out, cfa_stream_t stream) {
    if (n < 0 || !a || !b || !out) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_vec_mul, dim3(blocks_for(n)), dim3(BLOCK), 0, S(stream), n, a, b, out, nullptr);
    return launch_status();
}

int cfa_inv_diag(int n, const int* rp, const int* ci, const double* val, double eps, double* inv_diag,
                 cfa_stream_t stream) {
    if (n < 0 || !rp || !ci || !val || !inv_diag) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_inv_diag, dim3(blocks_for(n)), dim3(BLOCK), 0, S(stream), n, rp, ci, val, eps, inv_diag);
    return launch_status();
}

int cfa_amg_apply(const cfa_amg_t* amg, const double* r, double* z, cfa_stream_t stream) {
    if (!amg || !r || !z || !amg_valid(amg)) return CFA_E_ARG;
    amg_apply_enqueue(amg, r, z, nullptr, S(stream));
    return launch_status();
}

}  // extern "C"

namespace {

// PCG with the reductions folded into the consuming kernels (k_fp_*), two iterations per
// captured graph (the search direction and the r.z partials alternate between two buffers).
// Same algorithm, guards and stopping rules as the reference (src/cfa/linalg.cuh:183-256).
// The operator kernels run with 2 or 4 lanes per row (about four entries per lane).
int operator_lanes(int n, long long nnz) { return nnz <= 12LL * n ? 2 : 4; }

template <int TPR>
void launch_fp_operator_t(bool oss, int grid, hipStream_t s, const PcgArgs& a, int h) {
    if (oss) hipLaunchKernelGGL((k_fp_operator<TPR, true>), dim3(grid), dim3(BLOCK), 0, s, a, h);
    else hipLaunchKernelGGL((k_fp_operator<TPR, false>), dim3(grid), dim3(BLOCK), 0, s, a, h);
}
void launch_fp_operator(int tpr, bool oss, int grid, hipStream_t s, const PcgArgs& a, int h) {
    if (tpr == 2) launch_fp_operator_t<2>(oss, grid, s, a, h);
    else launch_fp_operator_t<4>(oss, grid, s, a, h);
}
void launch_apply_K_B(int tpr, int grid, hipStream_t s, int n, int nnz, const int* rp, const int* ci,
                      const double* val, const double* oss, const double* inv_m, const double* in, double* q,
                      double* gx, double* gy, const int* st) {
    if (tpr == 2)
        hipLaunchKernelGGL(k_apply_K_B<2>, dim3(grid), dim3(BLOCK), 0, s, n, nnz, rp, ci, val, oss, inv_m, in, q, gx, gy,
                           st);
    else
        hipLaunchKernelGGL(k_apply_K_B<4>, dim3(grid), dim3(BLOCK), 0, s, n, nnz, rp, ci, val, oss, inv_m, in, q, gx, gy,
                           st);
}
void launch_apply_BT_dot(int tpr, int grid, hipStream_t s, int n, int nnz, const int* rp, const int* ci,
                         const double* oss, const int* tag, OutletMode om, const double* gx, const double* gy,
                         const double* p, double* q, double* part, const int* st) {
    if (tpr == 2)
        hipLaunchKernelGGL(k_apply_BT_dot<2>, dim3(grid), dim3(BLOCK), 0, s, n, nnz, rp, ci, oss, tag, om, gx, gy, p,
                           q, part, st);
    else
        hipLaunchKernelGGL(k_apply_BT_dot<4>, dim3(grid), dim3(BLOCK), 0, s, n, nnz, rp, ci, oss, tag, om, gx, gy, p,
                           q, part, st);
}

// pressure-Dirichlet rows of the operator: node p_pin with the outlet Dirichlet condition, else outlet nodes
OutletMode outlet_of(const cfa_oss_op_t* oss) {
    return OutletMode{oss->outlet_dirichlet, oss->p_pin, 0.0};
}

int pcg_fused(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, int max_iter,
              double tol, const cfa_oss_op_t* oss, const cfa_amg_t* amg, cfa_krylov_ws_t* ws,
              cfa_solve_result_t* result, hipStream_t s) {
    PcgArgs a;
    std::memset(&a, 0, sizeof(a));   // the graph key compares the bytes, padding included
    a.n = n;
    a.max_iter = max_iter;
    a.rp = rp; a.ci = ci; a.val = val;
    a.r = ws->vec[0]; a.z = ws->vec[1]; a.p[0] = ws->vec[2]; a.q = ws->vec[3]; a.p[1] = ws->vec[5];
    double* inv_diag = ws->vec[4];
    a.inv_diag = inv_diag;
    a.x = x;
    // block partials: four segments of OP_BLOCKS doubles, then the prologue's (5 * 4096 of the
    // workspace's 8 * 4096)
    a.prz[0] = ws->part; a.prz[1] = ws->part + OP_BLOCKS; a.ppq = ws->part + 2 * OP_BLOCKS;
    a.prr = ws->part + 3 * OP_BLOCKS;
    double* scratch = ws->part + 4 * OP_BLOCKS;
    a.sc = ws->scal; a.st = state_words(ws); a.status = ws->status_dev;
    CFA_TRY(hipMemcpyAsync(&a.nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipStreamSynchronize(s));
    if (oss) {
        a.oss = oss->oss_val; a.inv_m = oss->inv_m; a.tag = oss->bc_tag; a.gx = oss->gx; a.gy = oss->gy;
        a.om = outlet_of(oss);
    }
    const int TPR = operator_lanes(n, a.nnz);
    const double init_sc[2] = {tol, -1.0};   // S_TOL, and S_REL "not computed"
    CFA_TRY(hipMemcpyAsync(a.sc + S_TOL, &init_sc[0], sizeof(double), hipMemcpyHostToDevice, s));
    CFA_TRY(hipMemcpyAsync(a.sc + S_REL, &init_sc[1], sizeof(double), hipMemcpyHostToDevice, s));
    ws->status_host[0] = 0;
    ws->status_host[1] = ST_RUNNING;
    const int nb = red_blocks(n);
    const int blocks = blocks_for(n);
    const int nb_op = std::min(blocks_for((long long)n * TPR), OP_BLOCKS);
    const int* st_r = a.st;

    // prologue (linalg.cuh:205-227): ||b||, x = 0 when trivial, r = b - A x, initial test, z = M^-1 r
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, s, n, b, b, scratch, nullptr);
    hipLaunchKernelGGL(k_pcg_init_bnorm, dim3(1), dim3(BLOCK), 0, s, nb, scratch, a.sc, a.st, a.status);
    hipLaunchKernelGGL(k_zero_if_trivial, dim3(blocks), dim3(BLOCK), 0, s, n, x, st_r);
    if (!amg) hipLaunchKernelGGL(k_inv_diag, dim3(blocks), dim3(BLOCK), 0, s, n, rp, ci, val, 1.0e-14, inv_diag);
    if (oss) {
        launch_apply_K_B(TPR, nb_op, s, n, a.nnz, rp, ci, val, a.oss, a.inv_m, x, a.q, a.gx, a.gy, st_r);
        launch_apply_BT_dot(TPR, nb_op, s, n, a.nnz, rp, ci, a.oss, a.tag, a.om, a.gx, a.gy, a.q, a.q, scratch, st_r);
    } else {
        csr_op_launch<OP_SPMV, double>(TPR, n, rp, ci, val, x, nullptr, nullptr, 0.0, a.q, st_r, s);
    }
    hipLaunchKernelGGL(k_pcg_residual, dim3(nb), dim3(BLOCK), 0, s, n, b, a.q, a.r, scratch, st_r);
    hipLaunchKernelGGL(k_pcg_init_rnorm, dim3(1), dim3(BLOCK), 0, s, nb, scratch, a.sc, a.st, a.status);
    auto apply_prec = [&](int h_next, bool check, hipStream_t q_s) {
        if (amg) {
            PrecHooks hk;
            if (check) {
                hk.prr = a.prr; hk.nb_rr = nb; hk.sc = a.sc; hk.st_w = a.st; hk.status = a.status;
            }
            hk.dot_part = a.prz[h_next];
            amg_apply_enqueue(amg, a.r, a.z, st_r, q_s, &hk);
            return hk.dot_blocks;
        }
        hipLaunchKernelGGL(k_fp_jacobi_prec, dim3(nb), dim3(BLOCK), 0, q_s, a, h_next, check ? 1 : 0);
        return nb;
    };
    a.nb_rz = apply_prec(0, false, s);
    a.nb_rr = nb;
    int rc = launch_status();
    if (rc) return rc;

    // iteration k with parity h (linalg.cuh:231-253)
    auto iteration = [&](int h, hipStream_t q_s) {
        launch_fp_operator(TPR, oss != nullptr, nb_op, q_s, a, h);
        if (oss)
            launch_apply_BT_dot(TPR, nb_op, q_s, n, a.nnz, rp, ci, a.oss, a.tag, a.om, a.gx, a.gy, a.p[h], a.q, a.ppq,
                                st_r);
        hipLaunchKernelGGL(k_fp_update, dim3(nb), dim3(BLOCK), 0, q_s, a, h, nb_op);
        (void)apply_prec(1 - h, true, q_s);
    };
    auto two_iterations = [&](hipStream_t q_s) {
        iteration(0, q_s);
        iteration(1, q_s);
    };
    hipGraphExec_t graph = nullptr;
    if (max_iter > 0) {
        // graph key: every pointer, size and constant the launches bake in
        KeyBuilder kb;
        kb.add(a);
        kb.add(oss != nullptr);
        kb.add(amg != nullptr);
        if (amg) {
            kb.add(amg->n_levels); kb.add(amg->coarse_n); kb.add(amg->coarse_inv); kb.add(amg->cb); kb.add(amg->cx);
            kb.add(amg->smooth_steps);
            if (amg->n_levels > 0) kb.add_bytes(amg->levels, sizeof(cfa_amg_level_t) * (size_t)amg->n_levels);
        }
        graph = iteration_graph(ws, internal_of(ws)->pcg, kb.k, two_iterations);
    }
    for (int g = 0; 2 * g < max_iter; ++g) {
        if (g >= 2) {
            wait_progress(ws, 2 * g - 3, s);
            if (ws->status_host[1] != ST_RUNNING) break;
        }
        if (graph) CFA_TRY(hipGraphLaunch(graph, s));
        else two_iterations(s);
    }
    rc = launch_status();
    if (rc) return rc;
    int fin[3];
    double rel;
    CFA_TRY(hipMemcpyAsync(fin, a.st, sizeof(fin), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipMemcpyAsync(&rel, a.sc + S_REL, sizeof(double), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipStreamSynchronize(s));
    if (fin[0] == ST_RUNNING || fin[0] == ST_CAPPED) {   // iteration cap
        result->iters[0] = max_iter;
        result->converged[0] = 0;
    } else {
        result->iters[0] = fin[1];
        result->converged[0] = fin[0] != ST_STOPPED;
    }
    result->rel_residual[0] = rel;
    return launch_status();
}

}  // namespace

extern "C" {

int cfa_to_f32(size_t n, const double* in, float* out, cfa_stream_t stream) {
    if ((n > 0 && (!in || !out))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    const int blocks = (int)std::min<size_t>((n + BLOCK - 1) / BLOCK, 4096);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks), dim3(BLOCK), 0, S(stream), n, in, out);
    return launch_status();
}

int cfa_amg_smooth(const cfa_amg_t* amg, int level, const double* x, const double* b, double* y,
                   cfa_stream_t stream) {
    if (!amg || !x || !b || !y || level < 0 || level >= amg->n_levels || !amg->levels) return CFA_E_ARG;
    const cfa_amg_level_t& L = amg->levels[level];
    if (L.s_sp && !sell_usable(L)) return CFA_E_ARG;
    level_op_launch<OP_JACOBI, double>(L, x, b, L.dinv, y, nullptr, S(stream));
    return launch_status();
}

int cfa_amg_smooth_f32(const cfa_amg_t* amg, int level, const float* x, const float* b, float* y,
                       cfa_stream_t stream) {
    if (!amg || !x || !b || !y || level < 0 || level >= amg->n_levels || !amg->levels) return CFA_E_ARG;
    const cfa_amg_level_t& L = amg->levels[level];
    if (!amg->use_f32 || !L.a_v32 || !L.dinv32 || (L.s_sp && !L.s_v32)) return CFA_E_NOTREADY;
    if (L.s_sp && !sell_usable(L)) return CFA_E_ARG;
    level_op_launch<OP_JACOBI, float>(L, x, b, L.dinv32, y, nullptr, S(stream));
    return launch_status();
}

size_t cfa_sell_ws_bytes(int rows) {
    if (rows < 0) return 0;
    const int slices = sell_positions(rows) / SELL_SLICE;
    const size_t scan = cfa_scan_counts_ws_bytes(slices);
    return scan ? ((size_t)slices * sizeof(int) + 255) / 256 * 256 + scan : 0;
}

int cfa_sell_plan(int rows, const int* rp, int* sell_row, int* sell_sp, int* info, void* ws, size_t ws_bytes,
                  cfa_stream_t stream) {
    if (rows < 0 || !info || !sell_sp) return CFA_E_ARG;
    if (rows > 0 && (!rp || !sell_row)) return CFA_E_ARG;
    const size_t need = cfa_sell_ws_bytes(rows);
    if (need == 0 || !ws || ws_bytes < need) return CFA_E_ARG;
    const hipStream_t s = S(stream);
    const int positions = sell_positions(rows), slices = positions / SELL_SLICE;
    int* cnt = static_cast<int*>(ws);
    const size_t cnt_bytes = ((size_t)slices * sizeof(int) + 255) / 256 * 256;
    CFA_TRY(hipMemsetAsync(info, 0, 2 * sizeof(int), s));
    if (rows > 0)
        hipLaunchKernelGGL(k_sell_plan, dim3(blocks_for(rows, SELL_WINDOW)), dim3(SELL_WINDOW), 0, s, rows, positions,
                           rp, sell_row, cnt, info);
    const int e = launch_status();
    if (e != CFA_OK) return e;
    return cfa_scan_counts(slices, cnt, sell_sp, info, static_cast<unsigned char*>(ws) + cnt_bytes, need - cnt_bytes,
                           stream);
}

int cfa_sell_fill(int rows, const int* rp, const int* ci, const double* val, const int* sell_row,
                  const int* sell_sp, int* sell_ci, double* sell_v, cfa_stream_t stream) {
    if (rows < 0) return CFA_E_ARG;
    if (rows == 0) return CFA_OK;
    if (!rp || !ci || !val || !sell_row || !sell_sp || !sell_ci || !sell_v) return CFA_E_ARG;
    const int positions = sell_positions(rows);
    hipLaunchKernelGGL(k_sell_fill, dim3(blocks_for(positions)), dim3(BLOCK), 0, S(stream), positions, rp, ci, val,
                       sell_row, sell_sp, sell_ci, sell_v);
    return launch_status();
}

int cfa_pcg_operator(int n, const int* rp, const int* ci, const double* val, const cfa_oss_op_t* oss,
                     const double* in, double* out, double* part, int stage, cfa_stream_t stream) {
    if (n < 0 || !rp || !ci || !val || !oss || !in || !out || !part || stage < 0 || stage > 2) return CFA_E_ARG;
    if (!oss->oss_val || !oss->inv_m || !oss->bc_tag || !oss->gx || !oss->gy) return CFA_E_ARG;
    if (oss->outlet_dirichlet && n > 0 && (oss->p_pin < 0 || oss->p_pin >= n)) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipStream_t s = S(stream);
    int nnz = 0;
    CFA_TRY(hipMemcpyAsync(&nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipStreamSynchronize(s));
    const int tpr = operator_lanes(n, nnz);
    const int blocks_tpr = std::min(blocks_for((long long)n * tpr), OP_BLOCKS);
    if (stage != 2)
        launch_apply_K_B(tpr, blocks_tpr, s, n, nnz, rp, ci, val, oss->oss_val, oss->inv_m, in, out, oss->gx, oss->gy,
                         nullptr);
    if (stage != 1)
        launch_apply_BT_dot(tpr, blocks_tpr, s, n, nnz, rp, ci, oss->oss_val, oss->bc_tag, outlet_of(oss), oss->gx,
                            oss->gy, in, out, part, nullptr);
    return launch_status();
}

int cfa_pcg(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, int max_iter,
            double tol, const cfa_oss_op_t* oss, cfa_op_cb extra_cb, void* extra_ctx, const cfa_amg_t* amg,
            cfa_op_cb prec_cb, void* prec_ctx, cfa_krylov_ws_t* ws, cfa_solve_result_t* result,
            cfa_stream_t stream) {
    if (n < 0 || !rp || !ci || !val || !b || !x || !ws || !result || ws->cap < n || !ws->status_host || max_iter < 0)
        return CFA_E_ARG;
    if (oss && (!oss->oss_val || !oss->inv_m || !oss->bc_tag || !oss->gx || !oss->gy)) return CFA_E_ARG;
    if (oss && oss->outlet_dirichlet && n > 0 && (oss->p_pin < 0 || oss->p_pin >= n)) return CFA_E_ARG;
    if (amg && !amg_valid(amg)) return CFA_E_ARG;
    hipStream_t s = S(stream);
    result->iters[0] = 0;
    result->converged[0] = 1;
    result->rel_residual[0] = -1.0;   // no residual computed (trivial or initial convergence)
    if (n == 0) return CFA_OK;
    if (!extra_cb && !prec_cb) return pcg_fused(n, rp, ci, val, b, x, max_iter, tol, oss, amg, ws, result, s);
    // solves with host callbacks: the iteration as separate kernels with one-workgroup finals
    double* r = ws->vec[0];
    double* z = ws->vec[1];
    double* p = ws->vec[2];
    double* q = ws->vec[3];
    double* inv_diag = ws->vec[4];
    double* part = ws->part;
    double* sc = ws->scal;
    int* st = state_words(ws);
    int* status = ws->status_dev;
    const int nb = red_blocks(n);
    const int blocks = blocks_for(n);
    constexpr int TPR = 4;
    const int blocks_tpr = std::min(blocks_for((long long)n * TPR), OP_BLOCKS);
    int nnz = 0;
    if (oss) {
        CFA_TRY(hipMemcpyAsync(&nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost, s));
        CFA_TRY(hipStreamSynchronize(s));
    }
    const double init_sc[2] = {tol, -1.0};   // S_TOL, and S_REL "not computed"
    CFA_TRY(hipMemcpyAsync(sc + S_TOL, &init_sc[0], sizeof(double), hipMemcpyHostToDevice, s));
    CFA_TRY(hipMemcpyAsync(sc + S_REL, &init_sc[1], sizeof(double), hipMemcpyHostToDevice, s));
    ws->status_host[0] = 0;
    ws->status_host[1] = ST_RUNNING;

    auto apply_op = [&](const double* in, double* out, hipStream_t q_s) {
        if (oss) {
            hipLaunchKernelGGL(k_apply_K_B<TPR>, dim3(blocks_tpr), dim3(BLOCK), 0, q_s, n, nnz, rp, ci, val,
                               oss->oss_val, oss->inv_m, in, out, oss->gx, oss->gy, st);
        } else {
            csr_op_launch<OP_SPMV, double>(TPR, n, rp, ci, val, in, nullptr, nullptr, 0.0, out, st, q_s);
            if (extra_cb) extra_cb(in, out, extra_ctx);
        }
    };
    // second stage of the fused operator (q -= B^T g) with the p.q partials, or a plain dot
    auto finish_op_dot = [&](const double* pv, double* qv, hipStream_t q_s) {
        if (oss) {
            hipLaunchKernelGGL(k_apply_BT_dot<TPR>, dim3(blocks_tpr), dim3(BLOCK), 0, q_s, n, nnz, rp, ci,
                               oss->oss_val, oss->bc_tag, outlet_of(oss), oss->gx, oss->gy, pv, qv, part, st);
            return blocks_tpr;
        }
        hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, q_s, n, pv, qv, part, st);
        return nb;
    };
    auto apply_prec = [&](const double* rv, double* zv, hipStream_t q_s) {
        if (amg) amg_apply_enqueue(amg, rv, zv, st, q_s);
        else if (prec_cb) prec_cb(rv, zv, prec_ctx);
        else hipLaunchKernelGGL(k_vec_mul, dim3(blocks), dim3(BLOCK), 0, q_s, n, inv_diag, rv, zv, st);
    };

    // prologue (src/cfa/linalg.cuh:205-227), all decisions on the device
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, s, n, b, b, part, nullptr);
    hipLaunchKernelGGL(k_pcg_init_bnorm, dim3(1), dim3(BLOCK), 0, s, nb, part, sc, st, status);
    hipLaunchKernelGGL(k_zero_if_trivial, dim3(blocks), dim3(BLOCK), 0, s, n, x, st);
    if (!amg && !prec_cb)
        hipLaunchKernelGGL(k_inv_diag, dim3(blocks), dim3(BLOCK), 0, s, n, rp, ci, val, 1.0e-14, inv_diag);
    apply_op(x, q, s);
    if (oss) (void)finish_op_dot(q, q, s);   // q -= B^T g (its dot is not used)
    hipLaunchKernelGGL(k_pcg_residual, dim3(nb), dim3(BLOCK), 0, s, n, b, q, r, part, st);
    hipLaunchKernelGGL(k_pcg_init_rnorm, dim3(1), dim3(BLOCK), 0, s, nb, part, sc, st, status);
    apply_prec(r, z, s);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(BLOCK), 0, s, n, z, p, st);
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, s, n, r, z, part, st);
    hipLaunchKernelGGL(k_pcg_final_beta, dim3(1), dim3(BLOCK), 0, s, nb, part, sc, st, status, 1);
    int rc = launch_status();
    if (rc) return rc;

    // one iteration (src/cfa/linalg.cuh:231-253)
    auto iteration = [&](hipStream_t q_s) {
        apply_op(p, q, q_s);
        const int npart = finish_op_dot(p, q, q_s);
        hipLaunchKernelGGL(k_pcg_final_alpha, dim3(1), dim3(BLOCK), 0, q_s, npart, part, sc, st, status);
        hipLaunchKernelGGL(k_pcg_update_xr, dim3(nb), dim3(BLOCK), 0, q_s, n, sc, p, q, x, r, part, st);
        hipLaunchKernelGGL(k_pcg_final_rnorm, dim3(1), dim3(BLOCK), 0, q_s, nb, part, sc, st, status);
        apply_prec(r, z, q_s);
        hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(BLOCK), 0, q_s, n, r, z, part, st);
        hipLaunchKernelGGL(k_pcg_final_beta, dim3(1), dim3(BLOCK), 0, q_s, nb, part, sc, st, status, 0);
        hipLaunchKernelGGL(k_pcg_p, dim3(blocks), dim3(BLOCK), 0, q_s, n, z, sc, p, st);
    };
    rc = run_iterations(ws, nullptr, max_iter, iteration, s);
    if (rc) return rc;
    int fin[3];
    double rel;
    CFA_TRY(hipMemcpyAsync(fin, st, sizeof(fin), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipMemcpyAsync(&rel, sc + S_REL, sizeof(double), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipStreamSynchronize(s));
    if (fin[0] == ST_RUNNING) {   // iteration cap
        result->iters[0] = max_iter;
        result->converged[0] = 0;
    } else {
        result->iters[0] = fin[1];
        result->converged[0] = fin[0] != ST_STOPPED;
    }
    result->rel_residual[0] = rel;
    return launch_status();
}

int cfa_ilu0_factor(const cfa_ilu_t* ilu, const double* val, cfa_stream_t stream) {
    if (!ilu || !val || !ilu->lu || !ilu->colour_start || !ilu->el_ptr || !ilu->el || !ilu->upd) return CFA_E_ARG;
    hipStream_t s = S(stream);
    if (ilu->n == 0) return CFA_OK;
    CFA_TRY(hipMemcpyAsync(ilu->lu, val, (size_t)ilu->nnz * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int c = 1; c < ilu->n_colours; ++c) {   // colour 0 has no L entries
        const int t0 = ilu->colour_start[c], m = ilu->colour_start[c + 1] - t0;
        if (m == 0) continue;
        hipLaunchKernelGGL(k_ilu_factor_rows, dim3(blocks_for(m, 128)), dim3(128), 0, s, m, t0, ilu->el_ptr, ilu->el,
                           ilu->upd, ilu->lu);
    }
    const int len = std::max(ilu->n, std::max(ilu->nl, ilu->nu));
    hipLaunchKernelGGL(k_ilu_gather, dim3(blocks_for(len)), dim3(BLOCK), 0, s, ilu->nl, ilu->nu, ilu->n, ilu->l_src,
                       ilu->u_src, ilu->d_src, ilu->lu, ilu->lv, ilu->uv, ilu->dinv, ilu->lv32, ilu->uv32, ilu->dinv32);
    if (ilu->s_lrow && ilu->s_urow && ilu->s_lsrc && ilu->s_usrc && ilu->s_lv && ilu->s_uv) {
        const int slots = std::max(std::max(ilu->s_lslots, ilu->s_uslots), 1);
        hipLaunchKernelGGL(k_ilu_sell_values, dim3(blocks_for(slots)), dim3(BLOCK), 0, s, ilu->s_lslots, ilu->s_uslots,
                           ilu->s_lsrc, ilu->s_usrc, ilu->lv, ilu->uv, ilu->lv32 ? ilu->lv32 : nullptr,
                           ilu->uv32 ? ilu->uv32 : nullptr, ilu->s_lv, ilu->s_uv, ilu->lv32 ? ilu->s_lv32 : nullptr,
                           ilu->uv32 ? ilu->s_uv32 : nullptr);
    }
    return launch_status();
}

int cfa_ilu0_apply(const cfa_ilu_t* ilu, const double* r, double* z, int nrhs, cfa_stream_t stream) {
    if (!ilu || !r || !z || nrhs < 1 || nrhs > 2 || !ilu->colour_start || !ilu->pos || !ilu->r_perm ||
        !ilu->z_perm)
        return CFA_E_ARG;
    if (ilu->n == 0) return CFA_OK;
    ilu_apply_any(ilu, 0, r, z, nrhs, nullptr, S(stream));
    return launch_status();
}

int cfa_bicgstab2(int n, const int* rp, const int* ci, const double* val, const double* b_u, const double* b_v,
                  double* x_u, double* x_v, int max_iter, double tol, const cfa_ilu_t* ilu, int use_f32,
                  cfa_krylov_ws_t* ws, cfa_solve_result_t* result, cfa_stream_t stream) {
    if (n < 0 || !rp || !ci || !val || !b_u || !b_v || !x_u || !x_v || !ilu || !ws || !result || ws->cap < n ||
        !ws->status_host || max_iter < 0 || !ilu->colour_start || ilu->n != n || !ilu->pos || !ilu->r_perm ||
        !ilu->z_perm)
        return CFA_E_ARG;
    hipStream_t s = S(stream);
    for (int c = 0; c < 2; ++c) { result->iters[c] = 0; result->converged[c] = 1; result->rel_residual[c] = 0.0; }
    if (n == 0) return CFA_OK;
    const size_t bytes = (size_t)n * sizeof(double);
    double* x = ws->vec[0];
    double* bb = ws->vec[1];
    double* r = ws->vec[2];
    double* r0 = ws->vec[3];
    double* p = ws->vec[4];
    double* v = ws->vec[5];
    double* sv = ws->vec[6];
    double* t = ws->vec[7];
    double* ph = ws->vec[8];
    double* sh = ws->vec[9];
    double* part_a = ws->part;                 // 6 * RED_BLOCKS each
    double* part_b = ws->part + 6 * RED_BLOCKS;
    double* part_c = ws->part + 12 * RED_BLOCKS;
    double* part_d = ws->part + 18 * RED_BLOCKS;
    double* sc = ws->scal;
    int* st = state_words(ws);
    int* status = ws->status_dev;
    CFA_TRY(hipMemcpyAsync(x, x_u, bytes, hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipMemcpyAsync(x + n, x_v, bytes, hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipMemcpyAsync(bb, b_u, bytes, hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipMemcpyAsync(bb + n, b_v, bytes, hipMemcpyDeviceToDevice, s));
    const double tol_v = tol;
    CFA_TRY(hipMemcpyAsync(sc + B_TOL, &tol_v, sizeof(double), hipMemcpyHostToDevice, s));
    ws->status_host[0] = 0;
    ws->status_host[1] = ST_RUNNING;

    // 2 lanes per row for rows of about 7 entries (4 for longer ones): with the partial sums
    // capped at RED_BLOCKS workgroups, 2 lanes also keep one row per group up to 131k rows
    int nnz = 0;
    CFA_TRY(hipMemcpyAsync(&nnz, rp + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipStreamSynchronize(s));
    const int TPR = operator_lanes(n, nnz);
    const int groups_per_block = BLOCK / TPR;
    const int nb_rows = std::min((n + groups_per_block - 1) / groups_per_block, RED_BLOCKS);
    const int nb2 = red_blocks(2LL * n);
    const int blocks2 = blocks_for(2LL * n);
    auto spmv2_dots = [&](const double* xin, double* yout, const double* w0, int mode, double* partials,
                          hipStream_t q_s) {
        if (TPR == 2)
            hipLaunchKernelGGL(k_spmv2_dots<2>, dim3(nb_rows), dim3(BLOCK), 0, q_s, n, rp, ci, val, xin, yout, w0, mode,
                               partials, st);
        else
            hipLaunchKernelGGL(k_spmv2_dots<4>, dim3(nb_rows), dim3(BLOCK), 0, q_s, n, rp, ci, val, xin, yout, w0, mode,
                               partials, st);
    };

    if (TPR == 2)
        hipLaunchKernelGGL(k_bicg_init<2>, dim3(nb_rows), dim3(BLOCK), 0, s, n, rp, ci, val, bb, x, r, r0, p, v, part_a);
    else
        hipLaunchKernelGGL(k_bicg_init<4>, dim3(nb_rows), dim3(BLOCK), 0, s, n, rp, ci, val, bb, x, r, r0, p, v, part_a);
    hipLaunchKernelGGL(k_bicg_final_init, dim3(1), dim3(BLOCK), 0, s, nb_rows, part_a, sc, st, status, max_iter);
    hipLaunchKernelGGL(k_bicg_pre, dim3(1), dim3(1), 0, s, sc, st, status);
    int rc = launch_status();
    if (rc) return rc;

    // one iteration for both components (src/cfa/linalg.cuh:700-775)
    auto iteration = [&](hipStream_t q_s) {
        hipLaunchKernelGGL(k_bicg_p, dim3(blocks2), dim3(BLOCK), 0, q_s, n, r, v, sc, p, st);
        ilu_apply_any(ilu, use_f32, p, ph, 2, st, q_s);
        spmv2_dots(ph, v, r0, 1, part_a, q_s);
        hipLaunchKernelGGL(k_bicg_s, dim3(nb2), dim3(BLOCK), 0, q_s, n, r, v, sc, sv, part_b, nb_rows, part_a, st,
                           status);
        ilu_apply_any(ilu, use_f32, sv, sh, 2, st, q_s);
        spmv2_dots(sh, t, sv, 2, part_c, q_s);
        hipLaunchKernelGGL(k_bicg_xr, dim3(nb2), dim3(BLOCK), 0, q_s, n, ph, sh, sv, t, r0, sc, x, r, part_d, nb2,
                           part_b, nb_rows, part_c, st);
        hipLaunchKernelGGL(k_bicg_final_post, dim3(1), dim3(BLOCK), 0, q_s, nb2, part_d, sc, st, status);
    };
    hipGraphExec_t graph = nullptr;
    if (max_iter > 0) {
        KeyBuilder kb;
        kb.add(n); kb.add(rp); kb.add(ci); kb.add(val); kb.add(use_f32); kb.add(TPR); kb.add(ilu->n_colours);
        kb.add_bytes(ilu->colour_start, sizeof(int) * (size_t)(ilu->n_colours + 1));
        kb.add(ilu->pos); kb.add(ilu->r_perm); kb.add(ilu->z_perm);
        kb.add(ilu->rows); kb.add(ilu->lp); kb.add(ilu->up); kb.add(ilu->lc); kb.add(ilu->uc); kb.add(ilu->lv);
        kb.add(ilu->uv); kb.add(ilu->dinv); kb.add(ilu->lv32); kb.add(ilu->uv32); kb.add(ilu->dinv32);
        for (int k = 0; k < 10; ++k) kb.add(ws->vec[k]);
        kb.add(ws->part); kb.add(sc); kb.add(status);
        graph = iteration_graph(ws, internal_of(ws)->bicg, kb.k, iteration);
    }
    rc = run_iterations(ws, graph, max_iter, iteration, s);
    if (rc) return rc;
    int fin[12];
    double rel[2];
    CFA_TRY(hipMemcpyAsync(fin, st, sizeof(fin), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipMemcpyAsync(rel, sc + B_REL, sizeof(rel), hipMemcpyDeviceToHost, s));
    CFA_TRY(hipMemcpyAsync(x_u, x, bytes, hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipMemcpyAsync(x_v, x + n, bytes, hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipStreamSynchronize(s));
    for (int c = 0; c < 2; ++c) {
        if (fin[2 + c]) {   // still active when the loop ended (cap reached without the in-kernel stop)
            result->iters[c] = max_iter;
            result->converged[c] = 0;
        } else {
            result->iters[c] = fin[4 + c];
            result->converged[c] = fin[6 + c];
        }
        result->rel_residual[c] = rel[c];
    }
    return launch_status();
}

int cfa_bicgstab_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
                        int max_iter, double tol, cfa_krylov_ws_t* ws, cfa_solve_result_t* result,
                        cfa_stream_t stream) {
    // Host-side scalar logic with one transfer per dot product, as the reference's one-component
    // solver (src/cfa/linalg.cuh:261-347); it is off the default path (--momentum-precond jacobi).
    if (n < 0 || !rp || !ci || !val || !b || !x || !ws || !result || ws->cap < n) return CFA_E_ARG;
    hipStream_t s = S(stream);
    result->iters[0] = 0;
    result->converged[0] = 1;
    result->rel_residual[0] = 0.0;
    if (n == 0) return CFA_OK;
    double* r = ws->vec[0];
    double* r0 = ws->vec[1];
    double* p = ws->vec[2];
    double* v = ws->vec[3];
    double* sv = ws->vec[4];
    double* t = ws->vec[5];
    double* ph = ws->vec[6];
    double* sh = ws->vec[7];
    double* dinv = ws->vec[8];
    const int blocks = blocks_for(n);
    auto dot = [&](const double* a, const double* bb) {
        double out = 0.0;
        cfa_dot(n, a, bb, &out, ws->part, stream);
        return out;
    };
    auto spmv = [&](const double* in, double* out) { cfa_spmv(n, rp, ci, val, in, out, 1, stream); };
    hipLaunchKernelGGL(k_inv_diag, dim3(blocks), dim3(BLOCK), 0, s, n, rp, ci, val, 1.0e-14, dinv);
    spmv(x, v);
    CFA_TRY(hipMemcpyAsync(r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, -1.0, v, r);
    CFA_TRY(hipMemcpyAsync(r0, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    CFA_TRY(hipMemcpyAsync(p, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(BLOCK), 0, s, n, v);
    const double rn0 = std::sqrt(dot(r, r));
    if (rn0 < 1.0e-13) return launch_status();
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    const double bd = 1.0e-30 * rn0 * rn0 + 1.0e-300;
    int iter = 0;
    bool done = false;
    for (; iter < max_iter; ++iter) {
        const double rho_new = dot(r0, r);
        if (std::fabs(rho_new) < bd) break;
        if (iter > 0) {
            const double beta = (rho_new / rho_old) * (alpha / (omega + 1.0e-30));
            hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, -omega, v, p);
            hipLaunchKernelGGL(k_xpay, dim3(blocks), dim3(BLOCK), 0, s, n, r, beta, p);
        }
        hipLaunchKernelGGL(k_vec_mul, dim3(blocks), dim3(BLOCK), 0, s, n, dinv, p, ph, nullptr);
        spmv(ph, v);
        const double r0v = dot(r0, v);
        if (std::fabs(r0v) < bd) break;
        alpha = rho_new / r0v;
        CFA_TRY(hipMemcpyAsync(sv, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, -alpha, v, sv);
        const double s_norm = std::sqrt(dot(sv, sv));
        if (s_norm < tol * rn0 || s_norm < 1.0e-11) {
            hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, alpha, ph, x);
            ++iter;
            done = true;
            break;
        }
        hipLaunchKernelGGL(k_vec_mul, dim3(blocks), dim3(BLOCK), 0, s, n, dinv, sv, sh, nullptr);
        spmv(sh, t);
        const double tt = dot(t, t);
        omega = (tt > bd) ? (dot(t, sv) / tt) : 0.0;
        hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, alpha, ph, x);
        hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, omega, sh, x);
        CFA_TRY(hipMemcpyAsync(r, sv, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(k_axpy, dim3(blocks), dim3(BLOCK), 0, s, n, -omega, t, r);
        const double r_norm = std::sqrt(dot(r, r));
        result->rel_residual[0] = r_norm / rn0;
        if (r_norm < tol * rn0 || r_norm < 1.0e-11) {
            ++iter;
            done = true;
            break;
        }
        rho_old = rho_new;
    }
    result->iters[0] = iter;
    result->converged[0] = done ? 1 : 0;
    CFA_TRY(hipStreamSynchronize(s));
    return launch_status();
}

}  // extern "C"
