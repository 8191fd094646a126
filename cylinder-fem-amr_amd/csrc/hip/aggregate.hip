// Greedy three-phase aggregation of the SA-AMG set-up on the device (the reference's
// smoothed_prolongator, src/cfa/linalg.cuh:1306-1356, runs it on the host in row order):
//   phase 1: in row order, a row none of whose strong neighbours is aggregated yet becomes a root
//            and takes its strong neighbours;
//   phase 2: a row left over joins the aggregate of its first aggregated strong neighbour;
//   phase 3: in row order, a row still left over becomes a root and takes its free strong
//            neighbours.
// The row-order greedy rules are lexicographically-first independent sets: a row is a phase-1
// root exactly when no lower root is in conflict with it (it claimed the row, is a strong
// neighbour of it, or claimed one of its strong neighbours), and a phase-3 root when no lower
// phase-3 root claimed it. Each decision needs the lower conflicting ones, so the decisions
// sweep through the rows as a front whose depth grows like the square root of the row count on
// these meshes. The lower conflicts of every row are listed in parallel first; then one workgroup
// decides the rows window by window (up to 8192 rows), with the window's conflict lists and
// decisions in LDS, in rounds that cost one barrier each. Aggregates are numbered in row order
// by prefix sums, as the greedy loop numbers them; the strength test and the P0 values
// 1/sqrt(size) are the host's floating-point operations (contraction off).
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "cfa_device.hpp"

using namespace cfa_dev;

#pragma clang fp contract(off)

namespace {

constexpr int FRONT_THREADS = 1024;
constexpr int FRONT_ROWS = 8;                                // window rows per thread
constexpr int FRONT_WINDOW = FRONT_THREADS * FRONT_ROWS;    // rows per window at most
constexpr int FRONT_LIST = 61440;                            // staged conflict entries (LDS, 120 KB)

int nblk(long long n) { return blocks_for(n > 0 ? n : 1); }

// conflict-list slot per row: four times the mean row length, 32 to 128 entries
int conflict_cap(int n, int nnz) {
    const long long c = 4LL * nnz / (n > 0 ? n : 1);
    return (int)std::min(128LL, std::max(32LL, (c + 7) / 8 * 8));
}

// strong entries per row (|a_ij| >= theta sqrt(|d_i d_j|), j != i, d = 1 / dinv) and the longest row
__global__ void k_strong_count(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ v, const double* __restrict__ dinv, double theta,
                               int* __restrict__ flag, int* __restrict__ cnt, int* __restrict__ info) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int len = 0;
    if (i < n) {
        const double di = 1.0 / dinv[i];
        int c = 0;
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int j = ci[a];
            const double dj = 1.0 / dinv[j];
            const int s = j != i && fabs(v[a]) >= theta * sqrt(fabs(di * dj));
            flag[a] = s;
            c += s;
        }
        cnt[i] = c;
        len = rp[i + 1] - rp[i];
    }
    // the longest row: one atomic per wave
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && len > 0) atomicMax(&info[2], len);
}

__global__ void k_fill_keys(int m, unsigned* __restrict__ keys, unsigned sentinel) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < m) keys[a] = sentinel;
}

// S (strong neighbours in row order) and the (column, row) pairs of the transpose
__global__ void k_strong_fill(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                              const int* __restrict__ flag, const int* __restrict__ s_rp, int* __restrict__ s_ci,
                              unsigned* __restrict__ keys, int* __restrict__ rows, int* __restrict__ t_cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = s_rp[i];
    for (int a = rp[i]; a < rp[i + 1]; ++a)
        if (flag[a]) {
            const int j = ci[a];
            s_ci[o] = j;
            keys[o] = (unsigned)j;
            rows[o] = i;
            ++o;
            atomicAdd(&t_cnt[j], 1);
        }
}

// Lower conflicts of each row, deduplicated, in a slot of `cap` entries per row: mode 1 (phase 1)
// the rows r < i that claimed i (T(i)), the strong neighbours r < i (S(i)) and the rows r < i that
// claimed a strong neighbour (T(j), j in S(i)); mode 3 (phase 3) T(i) only. len[i] = -1 marks a
// row with more than `cap` of them (its decisions walk the graph itself).
__global__ void k_lower_conflicts(int n, int mode, const int* __restrict__ s_rp, const int* __restrict__ s_ci,
                                  const int* __restrict__ t_rp, const int* __restrict__ t_ci, int cap,
                                  int* __restrict__ lc, int* __restrict__ len) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int* out = lc + (size_t)i * cap;
    int c = 0;
    bool full = false;
    auto add = [&](int r) {
        if (full) return;
        for (int k = 0; k < c; ++k)
            if (out[k] == r) return;
        if (c == cap) {
            full = true;
            return;
        }
        out[c++] = r;
    };
    for (int k = t_rp[i]; k < t_rp[i + 1]; ++k) {   // columns of T ascend
        const int r = t_ci[k];
        if (r >= i) break;
        add(r);
    }
    if (mode == 1)
        for (int k = s_rp[i]; k < s_rp[i + 1]; ++k) {
            const int j = s_ci[k];
            if (j < i) add(j);
            for (int q = t_rp[j]; q < t_rp[j + 1]; ++q) {
                const int r = t_ci[q];
                if (r >= i) break;
                add(r);
            }
        }
    len[i] = full ? -1 : c;
}

// exclusive scan over the 1024 threads of the workgroup; *total = the sum (s: 16 ints of LDS)
__device__ int block_scan_front(int v, int* s, int& total) {
    const int t = threadIdx.x, lane = t & (WAVE - 1), w = t / WAVE;
    int x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) s[w] = x;
    __syncthreads();
    if (w == 0) {
        int y = lane < FRONT_THREADS / WAVE ? s[lane] : 0;
#pragma unroll
        for (int o = 1; o < FRONT_THREADS / WAVE; o <<= 1) {
            const int z = __shfl_up(y, o, WAVE);
            if (lane >= o) y += z;
        }
        if (lane < FRONT_THREADS / WAVE) s[lane] = y;
    }
    __syncthreads();
    const int before = w > 0 ? s[w - 1] : 0;
    total = s[FRONT_THREADS / WAVE - 1];
    __syncthreads();
    return before + x - v;
}

// One workgroup deciding the greedy roots window by window. A row is a root exactly when none of
// its lower conflicts is a root; every row below the window is decided, so each window is
// decided completely by rounds in LDS: its rows' conflict lists are staged as window offsets
// (the conflicts below the window resolved from their final states while staging), and every
// round each undecided row either finds a lower root (not a root), or an undecided lower row (it
// waits on that row, rechecking only it until it is decided), or neither (a root). The lowest
// undecided row of the window always decides, so the rounds end. state (global): 0 undecided,
// 1 root, 2 not a root (rows outside the phase are 2 from the start).
__global__ __launch_bounds__(FRONT_THREADS) void k_front(int n, int mode, const int* __restrict__ s_rp,
                                                         const int* __restrict__ s_ci, const int* __restrict__ t_rp,
                                                         const int* __restrict__ t_ci, const int* __restrict__ lc,
                                                         const int* __restrict__ lc_len, int cap,
                                                         int* __restrict__ state, int* __restrict__ info) {
    __shared__ unsigned short s_list[FRONT_LIST];
    __shared__ unsigned char s_st[FRONT_WINDOW];
    __shared__ int s_scan[FRONT_THREADS / WAVE];
    __shared__ int s_flag[3];
    const int t = threadIdx.x;
    int rounds = 0;
    for (int lo = 0; lo < n;) {
        // the window: as wide as its staged lists allow
        int width = FRONT_WINDOW, hi = lo, base = 0, total = 0;
        while (true) {
            hi = min(n, lo + width);
            int mine = 0;
#pragma unroll
            for (int k = 0; k < FRONT_ROWS; ++k) {
                const int i = lo + t + k * FRONT_THREADS;
                if (i < hi) mine += max(lc_len[i], 0);
            }
            base = block_scan_front(mine, s_scan, total);
            if (total <= FRONT_LIST) break;
            width >>= 1;   // uniform: every thread saw the same total
        }
        const int wn = hi - lo;
        // staging
        int beg[FRONT_ROWS], cnt[FRONT_ROWS], wt[FRONT_ROWS];
        int off = base;
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            beg[k] = off;
            cnt[k] = 0;
            wt[k] = -1;
            if (loc >= wn) continue;
            const int i = lo + loc;
            const int raw = lc_len[i];
            unsigned char st0 = state[i] != 0 ? 2 : 0;
            if (raw < 0) {
                cnt[k] = -1;
            } else {
                const int* row = lc + (size_t)i * cap;
                int c = 0;
                for (int q = 0; q < raw; ++q) {
                    const int r = row[q];
                    if (r < lo) {
                        if (state[r] == 1) st0 = 2;
                    } else {
                        s_list[off + c++] = (unsigned short)(r - lo);
                    }
                }
                cnt[k] = c;
                off += raw;
            }
            s_st[loc] = st0;
        }
        if (t < 3) s_flag[t] = 0;
        __syncthreads();
        // rounds until every row of the window is decided
        for (int it = 0;; ++it) {
            if (t == 0) s_flag[(it + 1) % 3] = 0;
            bool pending = false;
#pragma unroll
            for (int k = 0; k < FRONT_ROWS; ++k) {
                const int loc = t + k * FRONT_THREADS;
                if (loc >= wn || s_st[loc] != 0) continue;
                if (wt[k] >= 0) {
                    const int sw = s_st[wt[k]];
                    if (sw == 0) {
                        pending = true;
                        continue;
                    }
                    if (sw == 1) {
                        s_st[loc] = 2;
                        continue;
                    }
                }
                int blocker = -1;
                bool lower_root = false;
                if (cnt[k] >= 0) {
                    for (int q = 0; q < cnt[k] && !lower_root; ++q) {
                        const int r = s_list[beg[k] + q];
                        const int sr = s_st[r];
                        if (sr == 1) lower_root = true;
                        else if (sr == 0 && blocker < 0) blocker = r;
                    }
                } else {   // a row with a long conflict list walks the graph
                    const int i = lo + loc;
                    auto see = [&](int r) {
                        const int sr = r < lo ? state[r] : s_st[r - lo];
                        if (sr == 1) lower_root = true;
                        else if (sr == 0 && blocker < 0) blocker = r - lo;
                    };
                    for (int q = t_rp[i]; q < t_rp[i + 1] && !lower_root; ++q) {
                        const int r = t_ci[q];
                        if (r >= i) break;
                        see(r);
                    }
                    if (mode == 1)
                        for (int q = s_rp[i]; q < s_rp[i + 1] && !lower_root; ++q) {
                            const int j = s_ci[q];
                            if (j < i) see(j);
                            for (int z = t_rp[j]; z < t_rp[j + 1] && !lower_root; ++z) {
                                const int r = t_ci[z];
                                if (r >= i) break;
                                see(r);
                            }
                        }
                }
                if (lower_root) {
                    s_st[loc] = 2;
                } else if (blocker >= 0) {
                    wt[k] = blocker;
                    pending = true;
                } else {
                    s_st[loc] = 1;
                }
            }
            if (pending) s_flag[it % 3] = 1;
            __syncthreads();
            ++rounds;
            if (s_flag[it % 3] == 0) break;
        }
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            if (loc < wn) state[lo + loc] = s_st[loc];
        }
        __syncthreads();   // LDS reused by the next window; its staging reads these states
        lo = hi;
    }
    if (t == 0) info[3] += rounds;
}

__global__ void k_init_state(int n, const int* __restrict__ agg, int* __restrict__ state) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    state[i] = (agg && agg[i] >= 0) ? 2 : 0;
}

__global__ void k_root_flags(int n, const int* __restrict__ state, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = state[i] == 1 ? 1 : 0;
}

// phase 1 result: roots numbered in row order, a claimed row takes the (only) root that claimed it
__global__ void k_phase1(int n, const int* __restrict__ state, const int* __restrict__ id, const int* __restrict__ t_rp,
                         const int* __restrict__ t_ci, int* __restrict__ agg1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = -1;
    if (state[i] == 1) {
        a = id[i];
    } else {
        for (int k = t_rp[i]; k < t_rp[i + 1]; ++k)
            if (state[t_ci[k]] == 1) { a = id[t_ci[k]]; break; }
    }
    agg1[i] = a;
}

// phase 2: the first aggregated strong neighbour in row order (phase-1 values)
__global__ void k_phase2(int n, const int* __restrict__ s_rp, const int* __restrict__ s_ci, const int* __restrict__ agg1,
                         int* __restrict__ agg2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = agg1[i];
    if (a < 0)
        for (int k = s_rp[i]; k < s_rp[i + 1]; ++k)
            if (agg1[s_ci[k]] >= 0) { a = agg1[s_ci[k]]; break; }
    agg2[i] = a;
}

// phase 3 result: roots numbered after the phase-1 aggregates, a claimed row takes the first
// (lowest) root that claimed it
__global__ void k_phase3(int n, const int* __restrict__ state, const int* __restrict__ id, const int* __restrict__ id_total1,
                         const int* __restrict__ t_rp, const int* __restrict__ t_ci, const int* __restrict__ agg2,
                         int* __restrict__ agg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = agg2[i];
    if (a < 0) {
        const int base = *id_total1;
        if (state[i] == 1) {
            a = base + id[i];
        } else {
            for (int k = t_rp[i]; k < t_rp[i + 1]; ++k) {
                const int r = t_ci[k];
                if (r >= i) break;
                if (state[r] == 1) { a = base + id[r]; break; }
            }
        }
    }
    agg[i] = a;
}

__global__ void k_agg_sizes(int n, const int* __restrict__ agg, int* __restrict__ size) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && agg[i] >= 0) atomicAdd(&size[agg[i]], 1);
}

__global__ void k_p0_values(int n, const int* __restrict__ agg, const int* __restrict__ size, double* __restrict__ p0v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p0v[i] = agg[i] >= 0 ? 1.0 / sqrt((double)size[agg[i]]) : 0.0;
}

__global__ void k_agg_total(const int* __restrict__ n1, const int* __restrict__ n3, int* __restrict__ info) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        info[0] = *n1 + *n3;
        info[1] = *n1;
    }
}

struct AggWs {
    int *flag, *cnt, *s_rp, *s_ci, *rows, *t_ci, *t_cnt, *t_rp, *state, *lc, *lc_len, *agg1, *agg2, *flags, *id, *id1,
        *size;
    int cap;
    unsigned *keys, *keys_sorted;
    void* tmp;
    size_t tmp_bytes;
};

size_t tmp_need(int n, int nnz) {
    size_t b = 0, need = 0;
    (void)rocprim::exclusive_scan(nullptr, b, static_cast<const int*>(nullptr), static_cast<int*>(nullptr), 0,
                                  (size_t)n + 1, rocprim::plus<int>());
    need = b;
    b = 0;
    (void)rocprim::radix_sort_pairs(nullptr, b, static_cast<const unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
                                    static_cast<const int*>(nullptr), static_cast<int*>(nullptr), (size_t)nnz, 0, 32);
    return (b > need ? b : need) + 256;
}

size_t layout(int n, int nnz, char* base, AggWs* w) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t N = (size_t)n + 2, M = (size_t)(nnz > 0 ? nnz : 1);
    AggWs t;
    t.flag = reinterpret_cast<int*>(take(M * 4));
    t.cnt = reinterpret_cast<int*>(take(N * 4));
    t.s_rp = reinterpret_cast<int*>(take(N * 4));
    t.s_ci = reinterpret_cast<int*>(take(M * 4));
    t.rows = reinterpret_cast<int*>(take(M * 4));
    t.t_ci = reinterpret_cast<int*>(take(M * 4));
    t.t_cnt = reinterpret_cast<int*>(take(N * 4));
    t.t_rp = reinterpret_cast<int*>(take(N * 4));
    t.state = reinterpret_cast<int*>(take(N * 4));
    t.cap = conflict_cap(n, nnz);
    t.lc = reinterpret_cast<int*>(take((size_t)n * t.cap * 4));
    t.lc_len = reinterpret_cast<int*>(take(N * 4));
    t.agg1 = reinterpret_cast<int*>(take(N * 4));
    t.agg2 = reinterpret_cast<int*>(take(N * 4));
    t.flags = reinterpret_cast<int*>(take(N * 4));
    t.id = reinterpret_cast<int*>(take(N * 4));
    t.id1 = reinterpret_cast<int*>(take(16));
    t.size = reinterpret_cast<int*>(take(N * 4));
    t.keys = reinterpret_cast<unsigned*>(take(M * 4));
    t.keys_sorted = reinterpret_cast<unsigned*>(take(M * 4));
    t.tmp_bytes = tmp_need(n, nnz);
    t.tmp = take(t.tmp_bytes);
    if (w) *w = t;
    return off;
}

}  // namespace

extern "C" {

size_t cfa_amg_aggregate_ws_bytes(int n, int nnz) {
    if (n <= 0 || nnz < 0) return 0;
    return layout(n, nnz, nullptr, nullptr);
}

int cfa_amg_aggregate(int n, int nnz, const int* rp, const int* ci, const double* val, const double* dinv,
                      double theta, void* ws, size_t ws_bytes, int* agg, double* p0v, int* info, cfa_stream_t stream) {
    if (n <= 0 || nnz <= 0 || !rp || !ci || !val || !dinv || !ws || !agg || !p0v || !info) return CFA_E_ARG;
    const size_t need = cfa_amg_aggregate_ws_bytes(n, nnz);
    if (ws_bytes < need) return CFA_E_ARG;
    AggWs w;
    layout(n, nnz, static_cast<char*>(ws), &w);
    const hipStream_t s = as_stream(stream);
    const dim3 gn(nblk(n)), b(BLOCK);
    hipError_t e = hipMemsetAsync(info, 0, 4 * sizeof(int), s);
    if (e == hipSuccess) e = hipMemsetAsync(w.t_cnt, 0, ((size_t)n + 1) * sizeof(int), s);
    if (e == hipSuccess) e = hipMemsetAsync(w.size, 0, (size_t)n * sizeof(int), s);
    if (e != hipSuccess) return (int)e;
    // strong graph S (row order) and its transpose (rows of each column ascending)
    hipLaunchKernelGGL(k_strong_count, gn, b, 0, s, n, rp, ci, val, dinv, theta, w.flag, w.cnt, info);
    hipLaunchKernelGGL(k_fill_keys, dim3(nblk(nnz)), b, 0, s, nnz, w.keys, (unsigned)n);
    hipLaunchKernelGGL(k_fill_keys, dim3(1), dim3(WAVE), 0, s, 1, reinterpret_cast<unsigned*>(w.cnt + n), 0u);
    size_t tb = w.tmp_bytes;
    e = rocprim::exclusive_scan(w.tmp, tb, w.cnt, w.s_rp, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_strong_fill, gn, b, 0, s, n, rp, ci, w.flag, w.s_rp, w.s_ci, w.keys, w.rows, w.t_cnt);
    tb = w.tmp_bytes;
    e = rocprim::radix_sort_pairs(w.tmp, tb, w.keys, w.keys_sorted, w.rows, w.t_ci, (size_t)nnz, 0, 32, s);
    if (e != hipSuccess) return (int)e;
    tb = w.tmp_bytes;
    e = rocprim::exclusive_scan(w.tmp, tb, w.t_cnt, w.t_rp, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    // phase 1
    hipLaunchKernelGGL(k_init_state, gn, b, 0, s, n, nullptr, w.state);
    hipLaunchKernelGGL(k_lower_conflicts, gn, b, 0, s, n, 1, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.cap, w.lc, w.lc_len);
    hipLaunchKernelGGL(k_front, dim3(1), dim3(FRONT_THREADS), 0, s, n, 1, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.lc,
                       w.lc_len, w.cap, w.state, info);
    hipLaunchKernelGGL(k_root_flags, gn, b, 0, s, n, w.state, w.flags);
    hipLaunchKernelGGL(k_fill_keys, dim3(1), dim3(WAVE), 0, s, 1, reinterpret_cast<unsigned*>(w.flags + n), 0u);
    tb = w.tmp_bytes;
    e = rocprim::exclusive_scan(w.tmp, tb, w.flags, w.id, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_phase1, gn, b, 0, s, n, w.state, w.id, w.t_rp, w.t_ci, w.agg1);
    e = hipMemcpyAsync(w.id1, w.id + n, sizeof(int), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int)e;
    // phase 2
    hipLaunchKernelGGL(k_phase2, gn, b, 0, s, n, w.s_rp, w.s_ci, w.agg1, w.agg2);
    // phase 3
    hipLaunchKernelGGL(k_init_state, gn, b, 0, s, n, w.agg2, w.state);
    hipLaunchKernelGGL(k_lower_conflicts, gn, b, 0, s, n, 3, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.cap, w.lc, w.lc_len);
    hipLaunchKernelGGL(k_front, dim3(1), dim3(FRONT_THREADS), 0, s, n, 3, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.lc,
                       w.lc_len, w.cap, w.state, info);
    hipLaunchKernelGGL(k_root_flags, gn, b, 0, s, n, w.state, w.flags);
    tb = w.tmp_bytes;
    e = rocprim::exclusive_scan(w.tmp, tb, w.flags, w.id, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_phase3, gn, b, 0, s, n, w.state, w.id, w.id1, w.t_rp, w.t_ci, w.agg2, agg);
    hipLaunchKernelGGL(k_agg_total, dim3(1), dim3(WAVE), 0, s, w.id1, w.id + n, info);
    // P0 values
    hipLaunchKernelGGL(k_agg_sizes, gn, b, 0, s, n, agg, w.size);
    hipLaunchKernelGGL(k_p0_values, gn, b, 0, s, n, agg, w.size, p0v);
    return launch_status();
}

}  // extern "C"
