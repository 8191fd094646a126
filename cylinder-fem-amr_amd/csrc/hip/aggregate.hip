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
constexpr int FRONT_LIST = 45056;                            // staged upper-list entries (LDS, 88 KB)

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
// row with more than `cap` of them (its decisions walk the graph itself). Rows outside the phase
// (state != 0) are never roots: they list nothing and are listed by no row.
__global__ void k_lower_conflicts(int n, int mode, const int* __restrict__ s_rp, const int* __restrict__ s_ci,
                                  const int* __restrict__ t_rp, const int* __restrict__ t_ci,
                                  const int* __restrict__ state, int cap, int* __restrict__ lc,
                                  int* __restrict__ len) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (state[i] != 0) {   // outside the phase: decided, and never a root, so never a conflict
        len[i] = 0;
        return;
    }
    int* out = lc + (size_t)i * cap;
    int c = 0;
    bool full = false;
    auto add = [&](int r) {
        if (full || state[r] != 0) return;
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

// appends `item` for the lanes with `want` to q at the LDS counter *qn (one atomic per wave)
__device__ __forceinline__ void wave_append(bool want, unsigned short item, unsigned short* q, int* qn) {
    const unsigned long long m = __ballot(want);
    if (!m) return;
    const int lane = threadIdx.x & (WAVE - 1), leader = __ffsll((long long)m) - 1;
    int b = 0;
    if (lane == leader) b = atomicAdd(qn, __popcll(m));
    b = __shfl(b, leader, WAVE);
    if (want) q[b + __popcll(m & ((1ull << lane) - 1))] = item;
}

// LDS word of a window row: bits 0-7 its lower conflicts in the window not yet decided, bit 8 a
// lower root seen, bit 9 a row with a long list (decides by walking the graph), bits 10-11 its
// decision once processed (1 root, 2 not a root), bits 16-31 the start of its upper list (the
// window rows that have it as a lower conflict; lists in row order, word wn holds the end).
constexpr unsigned W_FLAG = 1u << 8, W_LONG = 1u << 9, W_ROOT = 1u << 10, W_NOT = 2u << 10, W_DONE = 3u << 10;

// stage the lower conflicts of the window rows (vector loads of the row's slot): f(r) per entry
template <typename F>
__device__ __forceinline__ void for_lower(const int* __restrict__ lc, int cap, int i, int raw, F f) {
    const int4* row4 = reinterpret_cast<const int4*>(lc + (size_t)i * cap);
    for (int b = 0; b < raw; b += 16) {
        int4 v[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (b + 4 * c < raw) v[c] = row4[b / 4 + c];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e[4] = {v[c].x, v[c].y, v[c].z, v[c].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (b + 4 * c + j >= raw) break;
                f(e[j]);
            }
        }
    }
}

// One workgroup deciding the greedy roots window by window. A row is a root exactly when none of
// its lower conflicts is a root; every row below the window is decided when the window starts,
// so each window is decided completely inside LDS. Staging resolves the conflicts below the
// window from their final states and builds, for every row, the count of its undecided lower
// conflicts in the window and the list of window rows it is a lower conflict of. A row whose
// count reaches zero without a lower root (a root), or that meets a lower root (not a root), is
// queued by the update that decided it; each round processes the queue: the row's decision goes
// to `state`, and to its upper rows (a flag, or a decrement, by LDS atomics whose old value shows
// whether this update decided that row). Rows with more lower conflicts than their slot holds
// walk the graph every round until they decide, then join the queue. The lowest undecided row is
// always decidable, so every window ends. state (global): 0 undecided, 1 root, 2 not a root
// (rows outside the phase are 2 from the start).
__global__ __launch_bounds__(FRONT_THREADS) void k_front(int n, int mode, const int* __restrict__ s_rp,
                                                         const int* __restrict__ s_ci, const int* __restrict__ t_rp,
                                                         const int* __restrict__ t_ci, const int* __restrict__ lc,
                                                         const int* __restrict__ lc_len, int cap,
                                                         int* __restrict__ state, int* __restrict__ info) {
    __shared__ unsigned short s_up[FRONT_LIST];
    __shared__ unsigned s_word[FRONT_WINDOW + 1];
    __shared__ unsigned short s_q[2][FRONT_WINDOW];
    __shared__ int s_scan[FRONT_THREADS / WAVE];
    __shared__ int s_qn[3], s_pend[3];
    const int t = threadIdx.x;
    int rounds = 0;
    for (int lo = 0; lo < n;) {
        // the window: as many rows as their listed conflicts allow
        int width = FRONT_WINDOW, hi = lo, total = 0;
        while (true) {
            hi = min(n, lo + width);
            int mine = 0;
#pragma unroll
            for (int k = 0; k < FRONT_ROWS; ++k) {
                const int i = lo + t + k * FRONT_THREADS;
                if (i < hi) mine += max(lc_len[i], 0);
            }
            (void)block_scan_front(mine, s_scan, total);
            if (total <= FRONT_LIST) break;
            const int cut = (int)((long long)width * FRONT_LIST / total);   // uniform: every thread saw total
            width = max(1, min(cut, width - 1));
        }
        const int wn = hi - lo;
        // staging pass 1: counts of the undecided lower conflicts (registers), upper counts (LDS)
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            if (loc < wn) s_word[loc] = 0;
        }
        __syncthreads();
        unsigned own[FRONT_ROWS];
        unsigned longs = 0;   // row slots of this thread with a long list, undecided
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            own[k] = 0;
            if (loc >= wn) continue;
            const int i = lo + loc;
            const int raw = lc_len[i];
            unsigned w = 0;
            if (raw < 0) {
                w = W_LONG;
                longs |= 1u << k;
            } else {
                for_lower(lc, cap, i, raw, [&](int r) {
                    if (r < lo) {
                        if (state[r] == 1) w |= W_FLAG;
                    } else {
                        ++w;
                        atomicAdd(&s_word[r - lo], 1u);
                    }
                });
            }
            if (state[i] != 0) {   // outside the phase: not a root
                w |= W_FLAG;
                longs &= ~(1u << k);
            }
            own[k] = w;
        }
        __syncthreads();
        // pass 2: upper-list starts in row order (one scan per row slot), fill cursors in s_word
        int ucnt[FRONT_ROWS], ubeg[FRONT_ROWS];
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            ucnt[k] = loc < wn ? (int)s_word[loc] : 0;
        }
        int run = 0;
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            int tot = 0;
            ubeg[k] = run + block_scan_front(ucnt[k], s_scan, tot);
            run += tot;
        }
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            if (loc < wn) s_word[loc] = ubeg[k];
        }
        __syncthreads();
        // pass 3: the upper lists
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            if (loc >= wn) continue;
            const int i = lo + loc;
            const int raw = lc_len[i];
            if (raw > 0)
                for_lower(lc, cap, i, raw, [&](int r) {
                    if (r >= lo) s_up[atomicAdd(&s_word[r - lo], 1u)] = (unsigned short)loc;
                });
        }
        if (t < 3) {
            s_qn[t] = 0;
            s_pend[t] = 0;
        }
        __syncthreads();
        // row words and the first queue: rows decided by staging (a lower root below the window,
        // outside the phase, or no lower conflict in the window)
#pragma unroll
        for (int k = 0; k < FRONT_ROWS; ++k) {
            const int loc = t + k * FRONT_THREADS;
            bool want = false;
            unsigned short item = 0;
            if (loc < wn) {
                s_word[loc] = ((unsigned)ubeg[k] << 16) | own[k];
                if (own[k] & W_FLAG) {
                    want = true;
                    item = (unsigned short)loc;
                } else if (!(own[k] & W_LONG) && (own[k] & 0xffu) == 0) {
                    want = true;
                    item = (unsigned short)(loc | 0x8000);
                }
            }
            wave_append(want, item, s_q[0], &s_qn[0]);
        }
        if (t == 0) s_word[wn] = (unsigned)run << 16;
        __syncthreads();
        // rounds: queued rows publish their decision and update their upper rows; long rows walk
        for (int r = 0;; ++r) {
            const int cur = r % 3, nxt = (r + 1) % 3;
            const int qn = s_qn[cur];
            if (t == 0) {
                s_qn[(r + 2) % 3] = 0;
                s_pend[(r + 2) % 3] = 0;
            }
            const unsigned short* q = s_q[r & 1];
            unsigned short* qo = s_q[(r + 1) & 1];
            for (int base = 0; base < qn; base += FRONT_THREADS / 8) {
                const int qi = base + t / 8, sub = t & 7;
                int e = 0, e1 = 0;
                bool root = false;
                if (qi < qn) {
                    const unsigned item = q[qi];
                    const int loc = item & 0x7fff;
                    root = (item >> 15) != 0;
                    e = (int)(s_word[loc] >> 16) + sub;
                    e1 = (int)(s_word[loc + 1] >> 16);
                    if (sub == 0) {
                        state[lo + loc] = root ? 1 : 2;
                        atomicOr(&s_word[loc], root ? W_ROOT : W_NOT);
                    }
                }
                for (;; e += 8) {
                    const bool act = e < e1;
                    if (!__any(act)) break;
                    bool want = false;
                    unsigned short item = 0;
                    if (act) {
                        const int up = s_up[e];
                        if (root) {
                            want = !(atomicOr(&s_word[up], W_FLAG) & W_FLAG);
                            item = (unsigned short)up;
                        } else {
                            const unsigned old = atomicSub(&s_word[up], 1u);
                            want = (old & 0xffu) == 1 && !(old & W_FLAG);
                            item = (unsigned short)(up | 0x8000);
                        }
                    }
                    wave_append(want, item, qo, &s_qn[nxt]);
                }
            }
            // rows with long lists: a lower root decides "not a root", all lower rows decided "root"
            int pending = 0;
            if (longs) {
#pragma unroll
                for (int k = 0; k < FRONT_ROWS; ++k) {
                    if (!(longs & (1u << k))) continue;
                    const int loc = t + k * FRONT_THREADS, i = lo + loc;
                    bool lower_root = false, undecided = false;
                    auto see = [&](int x) {
                        const unsigned sx = x < lo ? (unsigned)state[x] << 10 : (s_word[x - lo] & W_DONE);
                        if (sx == W_ROOT) lower_root = true;
                        else if (sx == 0) undecided = true;
                    };
                    for (int a = t_rp[i]; a < t_rp[i + 1] && !lower_root; ++a) {
                        const int x = t_ci[a];
                        if (x >= i) break;
                        see(x);
                    }
                    if (mode == 1)
                        for (int a = s_rp[i]; a < s_rp[i + 1] && !lower_root; ++a) {
                            const int j = s_ci[a];
                            if (j < i) see(j);
                            for (int z = t_rp[j]; z < t_rp[j + 1] && !lower_root; ++z) {
                                const int x = t_ci[z];
                                if (x >= i) break;
                                see(x);
                            }
                        }
                    if (lower_root || !undecided) {
                        longs &= ~(1u << k);
                        qo[atomicAdd(&s_qn[nxt], 1)] = (unsigned short)(lower_root ? loc : (loc | 0x8000));
                    } else {
                        ++pending;
                    }
                }
            }
#pragma unroll
            for (int o = WAVE / 2; o > 0; o >>= 1) pending += __shfl_xor(pending, o, WAVE);
            if ((t & (WAVE - 1)) == 0 && pending) atomicAdd(&s_pend[nxt], pending);
            __syncthreads();
            ++rounds;
            if (s_qn[nxt] == 0 && s_pend[nxt] == 0) break;
        }
        __syncthreads();   // LDS reused by the next window, whose staging reads these states
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
    hipLaunchKernelGGL(k_lower_conflicts, gn, b, 0, s, n, 1, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.state, w.cap, w.lc,
                       w.lc_len);
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
    hipLaunchKernelGGL(k_lower_conflicts, gn, b, 0, s, n, 3, w.s_rp, w.s_ci, w.t_rp, w.t_ci, w.state, w.cap, w.lc,
                       w.lc_len);
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
