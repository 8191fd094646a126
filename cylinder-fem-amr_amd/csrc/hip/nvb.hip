// Newest-vertex-bisection forest on the device: marking with look-ahead, coarsening by vertex
// collapse, refinement rounds with conformity closure, Morton extraction and the P1 solution
// transfer. The reference runs all of this on the host (src/cfa/amr.cuh:58-996); here every phase
// is a set of data-parallel passes over the element or vertex slots:
//   * a pass of bisections is one list of leaves (ascending slot). The midpoints come from an
//     open-addressing table keyed by the bisected edge; an edge met by two leaves of the pass is
//     claimed by both and created by the earlier one (atomic minimum of the claim codes), so slots
//     are handed out by prefix sums in exactly the order a one-by-one loop over the list would use
//     (element and vertex slots recycled last-in first-out);
//   * conformity closure is iterated hanging-edge flags: every leaf with a midpoint on one of its
//     edges is bisected, pass after pass, until no leaf hangs;
//   * the element budget of the refinement marks is a sequential rule (an item is taken when the
//     predicted count stays within the budget); one workgroup applies it to chunks of the ordered
//     list with block prefix sums, taking each accepted prefix at once;
//   * look-ahead traces are one thread per leaf, the maximum reaching a leaf an atomic maximum on
//     the bit pattern of the non-negative indicator.
// Cylinder midpoints are logged for the host, which sets their arc positions with its own libm
// (atan2, cos, sin) before a later pass reads them, so coordinates match the reference's host
// code bit for bit. The look-ahead's distances use the device's hypot, sqrt and log2. The integer
// results (which leaves are bisected or merged, the slot each takes, the compact numbering)
// follow the reference's rules.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "cfa_device.hpp"

using namespace cfa_dev;

namespace {

constexpr unsigned long long H_EMPTY = ~0ull;
constexpr unsigned long long H_TOMB = ~0ull - 1ull;
constexpr int H_UNSET = 0x7fffffff;

// domain and cylinder (src/cfa/common.cuh:48-60)
constexpr double L_CH = 2.20, H_CH = 0.41, X_CYL = 0.20, Y_CYL = 0.20, R_CYL = 0.05;
constexpr double BUCKET = 0.01;   // look-ahead fallback grid cell (amr.cuh:565)

int nblk(long long n) { return blocks_for(n > 0 ? n : 1); }

__device__ __forceinline__ unsigned long long edge_key(int a, int b) {
    const unsigned lo = (unsigned)(a < b ? a : b), hi = (unsigned)(a < b ? b : a);
    return ((unsigned long long)lo << 32) | hi;
}

__device__ __forceinline__ unsigned hash_slot(unsigned long long k, unsigned mask) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (unsigned)k & mask;
}

// value stored for key k, or -1
__device__ __forceinline__ int hash_find(const cfa_forest_t& f, unsigned long long k) {
    const unsigned mask = (unsigned)f.hash_cap - 1u;
    unsigned h = hash_slot(k, mask);
    for (int probe = 0; probe < f.hash_cap; ++probe, h = (h + 1u) & mask) {
        const unsigned long long s = f.hkey[h];
        if (s == k) return f.hval[h];
        if (s == H_EMPTY) return -1;
    }
    return -1;
}

__device__ __forceinline__ bool is_leaf(const cfa_forest_t& f, int e) {
    return f.ealive[e] && f.ech[2 * e] < 0;
}

__device__ __forceinline__ double tri_area(const cfa_forest_t& f, int e) {
    const int a = f.ev[3 * e], b = f.ev[3 * e + 1], c = f.ev[3 * e + 2];
    return 0.5 * fabs((f.vx[b] - f.vx[a]) * (f.vy[c] - f.vy[a]) - (f.vx[c] - f.vx[a]) * (f.vy[b] - f.vy[a]));
}

// barycentric coordinates of (px, py) in element e, inside with a 1e-12 tolerance (amr.cuh:625-634)
__device__ __forceinline__ bool inside(const cfa_forest_t& f, int e, double px, double py, double l[3]) {
    const int a = f.ev[3 * e], b = f.ev[3 * e + 1], c = f.ev[3 * e + 2];
    const double x0 = f.vx[a], y0 = f.vy[a];
    const double d = (f.vx[b] - x0) * (f.vy[c] - y0) - (f.vx[c] - x0) * (f.vy[b] - y0);
    l[1] = ((px - x0) * (f.vy[c] - y0) - (f.vx[c] - x0) * (py - y0)) / d;
    l[2] = ((f.vx[b] - x0) * (py - y0) - (px - x0) * (f.vy[b] - y0)) / d;
    l[0] = 1.0 - l[1] - l[2];
    return l[1] >= -1.0e-12 && l[2] >= -1.0e-12 && l[1] + l[2] <= 1.0 + 1.0e-12;
}

// values of midpoint m of edge (a, b): P1 interpolation, exact Dirichlet data on boundary
// midpoints (amr.cuh:896-914)
__device__ __forceinline__ void midpoint_values(const cfa_forest_t& f, int m, int a, int b, double umax) {
    double u = 0.5 * (f.vu[a] + f.vu[b]), v = 0.5 * (f.vv[a] + f.vv[b]);
    const double p = 0.5 * (f.vp[a] + f.vp[b]);
    double uo = 0.5 * (f.vuo[a] + f.vuo[b]), vo = 0.5 * (f.vvo[a] + f.vvo[b]);
    const int t = f.vtag[m];
    if (t == BC_CYLINDER || t == BC_WALL) {
        u = v = uo = vo = 0.0;
    } else if (t == BC_INLET) {
        const double y = f.vy[m];
        u = uo = 4.0 * umax * y * (H_CH - y) / (H_CH * H_CH);
        v = vo = 0.0;
    }
    f.vu[m] = u;  f.vv[m] = v;  f.vp[m] = p;  f.vuo[m] = uo;  f.vvo[m] = vo;
}

__device__ __forceinline__ int take_slot(const int* freel, int nfree, int base, int j) {
    return j < nfree ? freel[nfree - 1 - j] : base + (j - nfree);
}

// ---- table -------------------------------------------------------------------------------------
__global__ void k_hash_clear(int cap, unsigned long long* __restrict__ hkey, int* __restrict__ hval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) {
        hkey[i] = H_EMPTY;
        hval[i] = H_UNSET;
    }
}

__global__ void k_hash_fill(cfa_forest_t f, int nv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nv || !f.valive[s] || f.vedge[s] == 0ull) return;
    const unsigned long long k = f.vedge[s];
    const unsigned mask = (unsigned)f.hash_cap - 1u;
    unsigned h = hash_slot(k, mask);
    for (int probe = 0; probe < f.hash_cap; ++probe, h = (h + 1u) & mask) {
        const unsigned long long old = atomicCAS(&f.hkey[h], H_EMPTY, k);
        if (old == H_EMPTY) {
            f.hval[h] = s;
            return;
        }
    }
}

// ---- leaf lists --------------------------------------------------------------------------------
__global__ void k_flag_leaves(cfa_forest_t f, int ne, int mode, const int* __restrict__ target, int* __restrict__ flags) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    int fl = is_leaf(f, e) ? 1 : 0;
    if (fl && mode == 1) fl = target[e] > f.egen[e] ? 1 : 0;
    if (fl && mode == 2) {
        const int a = f.ev[3 * e], b = f.ev[3 * e + 1], c = f.ev[3 * e + 2];
        fl = (hash_find(f, edge_key(a, b)) >= 0 || hash_find(f, edge_key(b, c)) >= 0 ||
              hash_find(f, edge_key(c, a)) >= 0) ? 1 : 0;
    }
    flags[e] = fl;
}

// ---- bisection pass ------------------------------------------------------------------------------
// Claims the midpoint of each listed leaf's refinement edge: mid[i] = the existing midpoint, or -1
// with the table slot in slot[i] and the claim code -(n - i) folded in by atomic minimum (the
// earliest leaf of the list wins).
__global__ void k_bisect_claim(cfa_forest_t f, int n, const int* __restrict__ list, int* __restrict__ mid,
                               int* __restrict__ slot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = list[i];
    if (e < 0 || e >= f.cnt[CFA_NVB_NE]) {   // not an element slot: reported, nothing touched
        atomicOr(&f.cnt[CFA_NVB_ERROR], 4);
        mid[i] = -3;
        slot[i] = 0;
        return;
    }
    const unsigned long long k = edge_key(f.ev[3 * e + 1], f.ev[3 * e + 2]);
    const int code = -(n - i);
    const unsigned mask = (unsigned)f.hash_cap - 1u;
    unsigned h = hash_slot(k, mask);
    for (int probe = 0; probe < f.hash_cap; ++probe, h = (h + 1u) & mask) {
        unsigned long long cur = f.hkey[h];
        if (cur == H_EMPTY) {
            cur = atomicCAS(&f.hkey[h], H_EMPTY, k);
            if (cur == H_EMPTY) cur = k;   // claimed this slot
        }
        if (cur != k) continue;
        const int v = f.hval[h];
        if (v >= 0 && v != H_UNSET) {   // a midpoint from before this pass
            mid[i] = v;
            slot[i] = h;
            return;
        }
        atomicMin(&f.hval[h], code);
        mid[i] = -1;
        slot[i] = h;
        return;
    }
    mid[i] = -2;   // table full: reported as an error
    slot[i] = 0;
}

__global__ void k_bisect_creators(cfa_forest_t f, int n, const int* __restrict__ mid, const int* __restrict__ slot,
                                  int* __restrict__ creator) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) creator[n] = 0;   // the scan's extra entry: rank[n] = midpoints created
    if (i >= n) return;
    creator[i] = (mid[i] == -1 && f.hval[slot[i]] == -(n - i)) ? 1 : 0;
    if (mid[i] == -2) atomicOr(&f.cnt[CFA_NVB_ERROR], 2);
    if (mid[i] < -1) creator[i] = 0;
}

// the creators make their midpoints: slot k-th of the pass's new vertices, coordinates (the
// cylinder arc for an edge between two cylinder vertices), tag, values, first parent, table entry
// and log entry
__global__ void k_bisect_vertices(cfa_forest_t f, int n, const int* __restrict__ list, const int* __restrict__ slot,
                                  const int* __restrict__ creator, const int* __restrict__ rank, double umax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !creator[i]) return;
    const int e = list[i];
    const int a = f.ev[3 * e + 1], b = f.ev[3 * e + 2];
    const int k = rank[i];
    const int m = take_slot(f.vfree, f.cnt[CFA_NVB_VFREE], f.cnt[CFA_NVB_NV], k);
    double mx = 0.5 * (f.vx[a] + f.vx[b]), my = 0.5 * (f.vy[a] + f.vy[b]);
    int tag = BC_INTERIOR;
    if (f.vtag[a] == BC_CYLINDER && f.vtag[b] == BC_CYLINDER) {
        const double th = atan2(my - Y_CYL, mx - X_CYL);
        mx = X_CYL + R_CYL * cos(th);
        my = Y_CYL + R_CYL * sin(th);
        tag = BC_CYLINDER;
    } else {   // straight boundaries by position (a point near the cylinder is interior here)
        const double eps = 1.0e-6;
        if (fabs(my) < eps || fabs(my - H_CH) < eps) tag = BC_WALL;
        else if (fabs(mx) < eps) tag = BC_INLET;
        else if (fabs(mx - L_CH) < eps) tag = BC_OUTLET;
    }
    f.vx[m] = mx;
    f.vy[m] = my;
    f.vtag[m] = tag;
    midpoint_values(f, m, a, b, umax);
    f.valive[m] = 1;
    f.vpar[2 * m] = e;
    f.vpar[2 * m + 1] = -1;
    f.vedge[m] = edge_key(a, b);
    f.hval[slot[i]] = m;
    const int lg = f.cnt[CFA_NVB_LOG] + k;
    f.log[3 * lg] = m;
    f.log[3 * lg + 1] = a;
    f.log[3 * lg + 2] = b;
    if (tag == BC_CYLINDER) {
        const int q = atomicAdd(&f.cnt[CFA_NVB_CYLINDER], 1);
        f.cyl_log[3 * q] = m;
        f.cyl_log[3 * q + 1] = a;
        f.cyl_log[3 * q + 2] = b;
    }
}

// children (m, c, a) and (m, b, c) in the element slots 2i and 2i + 1 of the pass; second parent of
// a midpoint met by another leaf
__global__ void k_bisect_split(cfa_forest_t f, int n, const int* __restrict__ list, const int* __restrict__ mid,
                               const int* __restrict__ slot, const int* __restrict__ creator, int inherit,
                               int* __restrict__ target) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || mid[i] < -1) return;
    const int e = list[i];
    const int c = f.ev[3 * e], a = f.ev[3 * e + 1], b = f.ev[3 * e + 2];
    const int m = mid[i] >= 0 ? mid[i] : f.hval[slot[i]];
    if (!creator[i] && atomicCAS(&f.vpar[2 * m + 1], -1, e) != -1) atomicOr(&f.cnt[CFA_NVB_ERROR], 1);
    const int nfree = f.cnt[CFA_NVB_EFREE], base = f.cnt[CFA_NVB_NE];
    const int c1 = take_slot(f.efree, nfree, base, 2 * i), c2 = take_slot(f.efree, nfree, base, 2 * i + 1);
    const int gen = f.egen[e];
    const int t = inherit ? target[e] : -1;
    f.ev[3 * c1] = m;  f.ev[3 * c1 + 1] = c;  f.ev[3 * c1 + 2] = a;
    f.ev[3 * c2] = m;  f.ev[3 * c2 + 1] = b;  f.ev[3 * c2 + 2] = c;
    f.egen[c1] = f.egen[c2] = gen + 1;
    f.eparent[c1] = f.eparent[c2] = e;
    f.ech[2 * c1] = f.ech[2 * c1 + 1] = -1;
    f.ech[2 * c2] = f.ech[2 * c2 + 1] = -1;
    f.ealive[c1] = f.ealive[c2] = 1;
    target[c1] = target[c2] = t;
    f.ech[2 * e] = c1;
    f.ech[2 * e + 1] = c2;
}

// slot counters after a pass of n bisections creating total new vertices
__global__ void k_bisect_counts(cfa_forest_t f, int n, const int* __restrict__ total) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int* c = f.cnt;
    const int ne_new = 2 * n, nv_new = *total;
    const int etake = ne_new < c[CFA_NVB_EFREE] ? ne_new : c[CFA_NVB_EFREE];
    c[CFA_NVB_NE] += ne_new - etake;
    c[CFA_NVB_EFREE] -= etake;
    const int vtake = nv_new < c[CFA_NVB_VFREE] ? nv_new : c[CFA_NVB_VFREE];
    c[CFA_NVB_NV] += nv_new - vtake;
    c[CFA_NVB_VFREE] -= vtake;
    c[CFA_NVB_LOG] += nv_new;
}

// ---- marking -----------------------------------------------------------------------------------
__device__ __forceinline__ double centroid_x(const cfa_forest_t& f, int e) {
    return (f.vx[f.ev[3 * e]] + f.vx[f.ev[3 * e + 1]] + f.vx[f.ev[3 * e + 2]]) / 3.0;
}
__device__ __forceinline__ double centroid_y(const cfa_forest_t& f, int e) {
    return (f.vy[f.ev[3 * e]] + f.vy[f.ev[3 * e + 1]] + f.vy[f.ev[3 * e + 2]]) / 3.0;
}

__global__ void k_normalise(cfa_forest_t f, cfa_nvb_mark_t m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    const int e = m.leaves[i];
    double w = 1.0;
    if (m.tol_radius > 0.0) {
        const double xc = centroid_x(f, e) - X_CYL, yc = centroid_y(f, e) - Y_CYL;
        const double d = sqrt(xc * xc + yc * yc);
        if (d > m.tol_radius) w = pow(d / m.tol_radius, m.tol_power);
    }
    if (m.tol_vort > 0.0) m.av[e] = m.eta[i] / (m.tol_vort * w);
    if (m.tol_interp > 0.0) m.ai[e] = m.eta[m.n_leaves + i] / (m.tol_interp * w);
}

// vertex -> leaf incidence of the leaves (compact ids)
__global__ void k_inc_count(cfa_forest_t f, int n_leaves, const int* __restrict__ leaves, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_leaves) return;
    const int e = leaves[i];
    for (int k = 0; k < 3; ++k) atomicAdd(&cnt[f.ev[3 * e + k]], 1);
}
__global__ void k_inc_fill(cfa_forest_t f, int n_leaves, const int* __restrict__ leaves, int* __restrict__ fill,
                           int* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_leaves) return;
    const int e = leaves[i];
    for (int k = 0; k < 3; ++k) list[atomicAdd(&fill[f.ev[3 * e + k]], 1)] = i;
}
// nbr[3 i + k]: the leaf sharing the edge opposite v[k] (the only other one on a conforming mesh)
__global__ void k_leaf_nbr(cfa_forest_t f, int n_leaves, const int* __restrict__ leaves, const int* __restrict__ ptr,
                           const int* __restrict__ list, int* __restrict__ nbr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_leaves) return;
    const int e = leaves[i];
    for (int k = 0; k < 3; ++k) {
        const int a = f.ev[3 * e + (k + 1) % 3], b = f.ev[3 * e + (k + 2) % 3];
        int found = -1;
        for (int s = ptr[a]; s < ptr[a + 1] && found < 0; ++s) {
            const int j = list[s];
            if (j == i) continue;
            const int* w = f.ev + 3 * leaves[j];
            if (w[0] == b || w[1] == b || w[2] == b) found = j;
        }
        nbr[3 * i + k] = found;
    }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The leaf containing (px, py) by the bucket-grid rule of the reference (amr.cuh:636-670): the first
// leaf, in leaf order, whose bounding-box cell range covers the point's cell and which contains
// the point; -1 when none does.
__device__ int locate_grid(const cfa_forest_t& f, int n_leaves, const int* __restrict__ leaves, double px, double py) {
    if (px < 0.0 || px > L_CH || py < 0.0 || py > H_CH) return -1;
    const int nx = (int)ceil(L_CH / BUCKET), ny = (int)ceil(H_CH / BUCKET);
    const int bi = clampi((int)(px / BUCKET), 0, nx - 1), bj = clampi((int)(py / BUCKET), 0, ny - 1);
    double l[3];
    for (int j = 0; j < n_leaves; ++j) {
        const int e = leaves[j];
        const double x0 = f.vx[f.ev[3 * e]], x1 = f.vx[f.ev[3 * e + 1]], x2 = f.vx[f.ev[3 * e + 2]];
        const double y0 = f.vy[f.ev[3 * e]], y1 = f.vy[f.ev[3 * e + 1]], y2 = f.vy[f.ev[3 * e + 2]];
        const double xmin = fmin(x0, fmin(x1, x2)), xmax = fmax(x0, fmax(x1, x2));
        const double ymin = fmin(y0, fmin(y1, y2)), ymax = fmax(y0, fmax(y1, y2));
        if (bi < clampi((int)(xmin / BUCKET), 0, nx - 1) || bi > clampi((int)(xmax / BUCKET), 0, nx - 1) ||
            bj < clampi((int)(ymin / BUCKET), 0, ny - 1) || bj > clampi((int)(ymax / BUCKET), 0, ny - 1))
            continue;
        if (inside(f, e, px, py, l)) return j;
    }
    return -1;
}

__device__ __forceinline__ double pow_q(double r, double q) {
    return q == 1.0 ? r : (q == 2.0 ? r * r : pow(r, q));
}

__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
    if (*p >= v) return;
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// one trace per leaf above tolerance (amr.cuh:711-759): centroid velocity for `horizon` seconds in
// at most 64 samples; each new leaf met receives src (h_T / h_K)^q
__global__ void k_lookahead(cfa_forest_t f, cfa_nvb_mark_t m, const double* __restrict__ src_v,
                            const double* __restrict__ src_i, const int* __restrict__ nbr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    const bool act_v = m.tol_vort > 0.0 && src_v[i] > 1.0, act_i = m.tol_interp > 0.0 && src_i[i] > 1.0;
    if (!act_v && !act_i) return;
    const double horizon = m.lookahead_time;
    const int e = m.leaves[i];
    const int a = f.ev[3 * e], b = f.ev[3 * e + 1], c = f.ev[3 * e + 2];
    const double xc = (f.vx[a] + f.vx[b] + f.vx[c]) / 3.0, yc = (f.vy[a] + f.vy[b] + f.vy[c]) / 3.0;
    const double uc = (f.vu[a] + f.vu[b] + f.vu[c]) / 3.0, vc = (f.vv[a] + f.vv[b] + f.vv[c]) / 3.0;
    const double hK = sqrt(2.0 * tri_area(f, e));
    const double dist = hypot(uc, vc) * horizon;
    if (dist < 0.5 * hK) return;
    const int ns = min(64, (int)ceil(dist / (0.5 * hK)));
    int last = i;
    double l[3];
    for (int s = 1; s <= ns; ++s) {
        const double fr = horizon * (double)s / ns;
        const double px = xc + uc * fr, py = yc + vc * fr;
        if (px < 0.0 || px > L_CH || py < 0.0 || py > H_CH) break;
        int t = last, found = -1;
        for (int hop = 0; hop < 256; ++hop) {
            if (inside(f, m.leaves[t], px, py, l)) { found = t; break; }
            const int k = (l[0] < l[1]) ? (l[0] < l[2] ? 0 : 2) : (l[1] < l[2] ? 1 : 2);
            const int nb = nbr[3 * t + k];
            if (nb < 0) break;
            t = nb;
        }
        if (found < 0) found = locate_grid(f, m.n_leaves, m.leaves, px, py);
        if (found < 0) break;
        if (found == last) continue;
        last = found;
        const int et = m.leaves[found];
        const double hT = sqrt(2.0 * tri_area(f, et));
        if (act_v) atomic_max_nonneg(&m.av[et], src_v[i] * pow_q(hT / hK, 1.0));
        if (act_i) atomic_max_nonneg(&m.ai[et], src_i[i] * pow_q(hT / hK, 2.0));
    }
}

__global__ void k_sources(cfa_nvb_mark_t m, double* __restrict__ src_v, double* __restrict__ src_i) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    src_v[i] = m.av[m.leaves[i]];
    src_i[i] = m.ai[m.leaves[i]];
}

// sort keys of the marking order: the score's bit pattern complemented (descending) for leaves
// above tolerance, all ones otherwise (after them)
__global__ void k_score_keys(cfa_nvb_mark_t m, unsigned long long* __restrict__ keys, int* __restrict__ vals,
                             int* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool above = false;
    if (i < m.n_leaves) {
        const int e = m.leaves[i];
        const double sc = fmax(m.av[e], m.ai[e]);
        above = sc > 1.0;
        keys[i] = above ? ~(unsigned long long)__double_as_longlong(sc) : ~0ull;
        vals[i] = i;
    }
    // one atomic per wave
    const unsigned long long mask = __ballot(above);
    if ((threadIdx.x & (WAVE - 1)) == 0 && mask) atomicAdd(count, (int)__popcll(mask));
}

// levels needed by the worst criterion (q = 1 vorticity, 2 interpolation): targets of the ordered
// leaves, as forest slots
__global__ void k_order_targets(cfa_forest_t f, cfa_nvb_mark_t m, const int* __restrict__ order,
                                const int* __restrict__ count, int* __restrict__ list, int* __restrict__ tgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const int e = m.leaves[order[i]];
    int levels = 1;
    if (m.av[e] > 1.0) levels = max(levels, (int)ceil(log2(m.av[e]) / 1.0));
    if (m.ai[e] > 1.0) levels = max(levels, (int)ceil(log2(m.ai[e]) / 2.0));
    list[i] = e;
    tgt[i] = f.egen[e] + 2 * levels;
}

// Sequential budget rule over an ordered list (amr.cuh:305-320 mark): item i (forest slot list[i],
// target generation tgt[i] capped at max_gen) is taken when tgt > gen and the predicted element
// count plus 2^(tgt - gen) - 1 stays within the budget. One workgroup of 256: per chunk, block prefix
// sums of the pending items' additions; the prefix that fits is taken at once, the first item that
// does not fit is dropped, and items that no longer fit whatever follows are dropped at once.
__global__ __launch_bounds__(256) void k_budget(cfa_forest_t f, cfa_nvb_mark_t m, const int* __restrict__ count,
                                                const int* __restrict__ list, const int* __restrict__ tgt,
                                                int tgt_const) {
    __shared__ long long scan[256];
    __shared__ long long s_pred;
    __shared__ int s_fail, s_any;
    const int tid = threadIdx.x;
    const int n = *count;
    const int max_gen = 2 * m.max_level;
    const long long cap = m.max_elems;
    if (tid == 0) s_pred = *m.predicted;
    __syncthreads();
    int taken = 0;
    for (int base = 0; base < n; base += 256) {
        const int i = base + tid;
        int e = -1, t = 0;
        long long add = 0;
        if (i < n) {
            e = list[i];
            const int g = f.egen[e];
            t = min(max_gen, tgt ? tgt[i] : tgt_const);
            if (t > g) add = (1ll << (t - g)) - 1;
        }
        bool pending = add > 0;
        while (true) {
            if (tid == 0) s_any = 0;
            __syncthreads();
            if (pending) s_any = 1;
            __syncthreads();
            if (!s_any) break;
            // inclusive scan of the pending additions in lane order
            scan[tid] = pending ? add : 0;
            __syncthreads();
            for (int o = 1; o < 256; o <<= 1) {
                const long long v = tid >= o ? scan[tid - o] : 0;
                __syncthreads();
                scan[tid] += v;
                __syncthreads();
            }
            const long long pred = s_pred;
            if (tid == 0) s_fail = 256;
            __syncthreads();
            if (pending && pred + scan[tid] > cap) atomicMin(&s_fail, tid);
            __syncthreads();
            const int fail = s_fail;
            const long long used = fail > 0 ? scan[fail - 1] : 0;
            if (pending && tid < fail) {
                m.refine[e] = 1;
                m.target[e] = t;
                ++taken;
                pending = false;
            }
            if (tid == fail) pending = false;
            __syncthreads();
            if (tid == 0) s_pred = pred + used;
            __syncthreads();
            if (pending && s_pred + add > cap) pending = false;
        }
    }
    __syncthreads();
    if (taken) atomicAdd(&f.cnt[CFA_NVB_MARKED_REFINE], taken);
    if (tid == 0) *m.predicted = s_pred;
}

__device__ __forceinline__ int planned_level(const cfa_forest_t& f, const cfa_nvb_mark_t& m, int e) {
    return max(f.egen[e], m.target[e]) / 2;
}
__device__ __forceinline__ int finer_neighbours(const cfa_forest_t& f, const cfa_nvb_mark_t& m,
                                                const int* __restrict__ nbr, int i) {
    const int lvl = planned_level(f, m, m.leaves[i]);
    int c = 0;
    for (int k = 0; k < 3; ++k) {
        const int nb = nbr[3 * i + k];
        if (nb >= 0 && planned_level(f, m, m.leaves[nb]) > lvl) ++c;
    }
    return c;
}

// candidate flags over the leaves (leaf order): 0 minimum level, 1 hole filling, 2 buffer layer
__global__ void k_candidates(cfa_forest_t f, cfa_nvb_mark_t m, int kind, const int* __restrict__ nbr,
                             int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    const int e = m.leaves[i];
    const int gen = f.egen[e], max_gen = 2 * m.max_level;
    int fl = 0;
    if (!m.refine[e]) {
        if (kind == 0) {
            fl = gen < 2 * m.min_level;
        } else if (kind == 1) {
            fl = gen < max_gen && fmax(m.av[e], m.ai[e]) > m.coarsen_safety && finer_neighbours(f, m, nbr, i) >= 2;
        } else {
            const int* w = f.ev + 3 * e;
            fl = gen < max_gen && (m.vflag[w[0]] || m.vflag[w[1]] || m.vflag[w[2]]);
        }
    }
    flags[i] = fl;
}

__global__ void k_leaf_slots(cfa_nvb_mark_t m, const int* __restrict__ count, int* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *count) list[i] = m.leaves[list[i]];
}

__global__ void k_plus_two(cfa_forest_t f, const int* __restrict__ count, const int* __restrict__ list,
                           int* __restrict__ tgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *count) tgt[i] = f.egen[list[i]] + 2;
}

__global__ void k_vflag_marked(cfa_forest_t f, cfa_nvb_mark_t m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    const int e = m.leaves[i];
    if (!m.refine[e]) return;
    for (int k = 0; k < 3; ++k) m.vflag[f.ev[3 * e + k]] = 1;
}

// coarsening flags (amr.cuh:374-389); flags over the leaves for the leaf-order list
__global__ void k_coarsen_flags(cfa_forest_t f, cfa_nvb_mark_t m, const int* __restrict__ nbr, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_leaves) return;
    const int e = m.leaves[i];
    const int* w = f.ev + 3 * e;
    int fl = !m.refine[e] && f.egen[e] > 2 * m.min_level && !(m.vflag[w[0]] || m.vflag[w[1]] || m.vflag[w[2]]);
    if (fl && m.keep_radius > 0.0) {
        const double xc = centroid_x(f, e) - X_CYL, yc = centroid_y(f, e) - Y_CYL;
        if (xc * xc + yc * yc < m.keep_radius * m.keep_radius) fl = 0;
    }
    if (fl) fl = fmax(1.4142135623730951 * m.av[e], 2.0 * m.ai[e]) < m.coarsen_safety &&
                 finer_neighbours(f, m, nbr, i) < 2;
    if (fl) m.coarsen[e] = 1;
    flags[i] = fl;
}

// ---- coarsening sweep --------------------------------------------------------------------------
__device__ __forceinline__ double parent_score(const cfa_nvb_mark_t& m, int e) {
    return fmax(1.4142135623730951 * m.av[e], 2.0 * m.ai[e]);
}

__global__ void k_collapse_check(cfa_forest_t f, cfa_nvb_mark_t m, int n_max, const int* __restrict__ count,
                                 int* __restrict__ vcand) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_max || i >= *count) return;
    const int e = m.clist[i];
    if (!is_leaf(f, e) || f.egen[e] == 0) return;
    const int v = f.ev[3 * e];
    bool ok = false;
    for (int s = 0; s < 2; ++s) {
        const int pe = f.vpar[2 * v + s];
        if (pe < 0) continue;
        const int c0 = f.ech[2 * pe], c1 = f.ech[2 * pe + 1];
        ok = f.ealive[pe] && c0 >= 0 && is_leaf(f, c0) && is_leaf(f, c1) && m.coarsen[c0] && m.coarsen[c1] &&
             fmax(parent_score(m, c0), parent_score(m, c1)) < m.coarsen_safety;
        if (!ok) break;
    }
    if (ok) vcand[v] = 1;
}

__global__ void k_collapse_sizes(cfa_forest_t f, const int* __restrict__ go, const int* __restrict__ count, int n_max,
                                 int* __restrict__ sizes) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_max) return;
    int s = 0;
    if (k < *count) {
        const int v = go[k];
        s = (f.vpar[2 * v] >= 0 ? 1 : 0) + (f.vpar[2 * v + 1] >= 0 ? 1 : 0);
    }
    sizes[k] = s;
}

__global__ void k_collapse(cfa_forest_t f, cfa_nvb_mark_t m, const int* __restrict__ go, const int* __restrict__ count,
                           const int* __restrict__ off) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = *count;
    if (k >= g) return;
    const int v = go[k];
    const int nefree = f.cnt[CFA_NVB_EFREE], nvfree = f.cnt[CFA_NVB_VFREE];
    int q = off[k];   // merged parents before this vertex
    for (int s = 0; s < 2; ++s) {
        const int pe = f.vpar[2 * v + s];
        if (pe < 0) continue;
        const int c0 = f.ech[2 * pe], c1 = f.ech[2 * pe + 1];
        const double mv = 1.4142135623730951 * fmax(m.av[c0], m.av[c1]);
        const double mi = 2.0 * fmax(m.ai[c0], m.ai[c1]);
        for (int c : {c0, c1}) {
            f.ealive[c] = 0;
            f.ech[2 * c] = f.ech[2 * c + 1] = -1;
            f.eparent[c] = -1;
            f.egen[c] = 0;
        }
        f.efree[nefree + 2 * q] = c0;
        f.efree[nefree + 2 * q + 1] = c1;
        f.ech[2 * pe] = f.ech[2 * pe + 1] = -1;
        m.coarsen[pe] = 1;
        m.av[pe] = mv;
        m.ai[pe] = mi;
        m.clist[q] = pe;
        ++q;
    }
    // erase the bisected edge from the table
    const unsigned long long key = f.vedge[v];
    const unsigned mask = (unsigned)f.hash_cap - 1u;
    unsigned h = hash_slot(key, mask);
    for (int probe = 0; probe < f.hash_cap; ++probe, h = (h + 1u) & mask) {
        const unsigned long long s = f.hkey[h];
        if (s == key) { f.hkey[h] = H_TOMB; break; }
        if (s == H_EMPTY) break;
    }
    f.valive[v] = 0;
    f.vpar[2 * v] = f.vpar[2 * v + 1] = -1;
    f.vfree[nvfree + k] = v;
}

__global__ void k_collapse_counts(cfa_forest_t f, const int* __restrict__ count, const int* __restrict__ off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int g = *count;
    const int merged = off[g];
    f.cnt[CFA_NVB_EFREE] += 2 * merged;
    f.cnt[CFA_NVB_VFREE] += g;
    f.cnt[CFA_NVB_REMOVED] += g;
    f.cnt[CFA_NVB_CANDIDATES] = g;
    f.cnt[CFA_NVB_LIST] = merged;
}

// ---- extraction --------------------------------------------------------------------------------
__device__ __forceinline__ unsigned spread16(unsigned v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__device__ __forceinline__ unsigned quant16(double s) {
    return (unsigned)fmin(65535.0, fmax(0.0, s * 65535.0));
}

__global__ void k_morton(cfa_forest_t f, int n, const int* __restrict__ leaves, unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = leaves[i];
    keys[i] = spread16(quant16(centroid_x(f, e) / L_CH)) | (spread16(quant16(centroid_y(f, e) / H_CH)) << 1);
}

__global__ void k_fill_int(int n, int* __restrict__ p, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_first_use(cfa_forest_t f, int n, const int* __restrict__ leaf, int* __restrict__ first) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 3 * n) return;
    atomicMin(&first[f.ev[3 * leaf[p / 3] + p % 3]], p);
}

__global__ void k_first_flags(cfa_forest_t f, int n, const int* __restrict__ leaf, const int* __restrict__ first,
                              int* __restrict__ flags) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 3 * n) return;
    flags[p] = first[f.ev[3 * leaf[p / 3] + p % 3]] == p ? 1 : 0;
}

// node j: its vertex slot, position and values; compact[slot] = j
__global__ void k_nodes(cfa_forest_t f, const int* __restrict__ count, int max_n, const int* __restrict__ leaf,
                        const int* __restrict__ pos, int* __restrict__ node, int* __restrict__ compact, double* x,
                        double* y, int* tag, double* values) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int nn = *count;
    if (j >= nn || j >= max_n) return;
    const int p = pos[j];
    const int v = f.ev[3 * leaf[p / 3] + p % 3];
    node[j] = v;
    compact[v] = j;
    x[j] = f.vx[v];
    y[j] = f.vy[v];
    tag[j] = f.vtag[v];
    values[j] = f.vu[v];
    values[(size_t)nn + j] = f.vv[v];
    values[2 * (size_t)nn + j] = f.vp[v];
    values[3 * (size_t)nn + j] = f.vuo[v];
    values[4 * (size_t)nn + j] = f.vvo[v];
}

__global__ void k_elements(cfa_forest_t f, int n, const int* __restrict__ leaf, const int* __restrict__ compact,
                           int* __restrict__ elem_v, int* __restrict__ gen) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = leaf[i];
    for (int k = 0; k < 3; ++k) elem_v[3 * i + k] = compact[f.ev[3 * e + k]];
    gen[i] = f.egen[e];
}

__global__ void k_uniform_targets(cfa_forest_t f, int n, const int* __restrict__ leaves, int levels,
                                  int* __restrict__ target) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) target[leaves[i]] = f.egen[leaves[i]] + 2 * levels;
}

__global__ void k_copy_count(const int* __restrict__ src, int* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

// ---- transfer ----------------------------------------------------------------------------------
__global__ void k_scatter(cfa_forest_t f, int n, const int* __restrict__ nodes, const double* __restrict__ val) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int s = nodes[j];
    f.vu[s] = val[j];
    f.vv[s] = val[(size_t)n + j];
    f.vp[s] = val[2 * (size_t)n + j];
    f.vuo[s] = val[3 * (size_t)n + j];
    f.vvo[s] = val[4 * (size_t)n + j];
}

__global__ void k_set_xy(cfa_forest_t f, int n, const int* __restrict__ slots, const double* __restrict__ xy) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    f.vx[slots[j]] = xy[2 * j];
    f.vy[slots[j]] = xy[2 * j + 1];
}

__global__ void k_gather(cfa_forest_t f, int n, const int* __restrict__ nodes, double* __restrict__ val) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int s = nodes[j];
    val[j] = f.vu[s];
    val[(size_t)n + j] = f.vv[s];
    val[2 * (size_t)n + j] = f.vp[s];
    val[3 * (size_t)n + j] = f.vuo[s];
    val[4 * (size_t)n + j] = f.vvo[s];
}

__global__ void k_replay(cfa_forest_t f, int first, int last, double umax) {
    const int q = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= last) return;
    midpoint_values(f, f.log[3 * q], f.log[3 * q + 1], f.log[3 * q + 2], umax);
}

// ---- rocPRIM helpers ---------------------------------------------------------------------------
hipError_t select_flagged(void* tmp, size_t tmp_bytes, const int* flags, int* out, int* count, int n, hipStream_t s) {
    size_t tb = tmp_bytes;
    return rocprim::select(tmp, tb, rocprim::counting_iterator<int>(0), flags, out, count, (size_t)n, s);
}

hipError_t scan_ints(void* tmp, size_t tmp_bytes, const int* in, int* out, int n, hipStream_t s) {
    size_t tb = tmp_bytes;
    return rocprim::exclusive_scan(tmp, tb, in, out, 0, (size_t)n, rocprim::plus<int>(), s);
}

size_t align8(size_t ints) { return (ints + 1) & ~(size_t)1; }

#define GPU_TRY(call)                                \
    do {                                             \
        const hipError_t e_ = (call);                \
        if (e_ != hipSuccess) return (int)e_;        \
    } while (0)

}  // namespace

extern "C" {

size_t cfa_nvb_tmp_bytes(int n) {
    if (n < 1) n = 1;
    size_t need = 0, b = 0;
    (void)rocprim::select(nullptr, b, rocprim::counting_iterator<int>(0), static_cast<const int*>(nullptr),
                    static_cast<int*>(nullptr), static_cast<int*>(nullptr), (size_t)n);
    need = b > need ? b : need;
    b = 0;
    (void)rocprim::exclusive_scan(nullptr, b, static_cast<const int*>(nullptr), static_cast<int*>(nullptr), 0,
                            (size_t)n + 1, rocprim::plus<int>());
    need = b > need ? b : need;
    b = 0;
    (void)rocprim::radix_sort_pairs(nullptr, b, static_cast<const unsigned long long*>(nullptr),
                              static_cast<unsigned long long*>(nullptr), static_cast<const int*>(nullptr),
                              static_cast<int*>(nullptr), (size_t)n, 0, 64);
    need = b > need ? b : need;
    b = 0;
    (void)rocprim::radix_sort_pairs(nullptr, b, static_cast<const unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
                              static_cast<const int*>(nullptr), static_cast<int*>(nullptr), (size_t)n, 0, 32);
    need = b > need ? b : need;
    return need + 256;
}

// layout (ints): bisect 4 n + 2; mark: 10 L + 3 L + 3 nv + 4; coarsen: 4 nv + 4; extract:
// 12 L + 2 nv + 8 (8-byte aligned pieces)
size_t cfa_nvb_scratch_ints(int n_elem, int n_vert, int n_leaves) {
    const size_t ne = n_elem > 0 ? n_elem : 1, nv = n_vert > 0 ? n_vert : 1, L = n_leaves > 0 ? n_leaves : 1;
    size_t need = 4 * ne + 8;
    const size_t mark = align8(2 * L) * 2 + align8(nv + 2) * 2 + align8(nv) + align8(3 * L) + align8(3 * L) +
                        2 * align8(2 * L) + 2 * align8(L) + align8(L) + align8(L) + align8(L) + 8;
    const size_t coarsen = 4 * align8(nv + 2) + 8;
    const size_t extract = align8(L) * 3 + align8(L) + align8(nv) * 2 + align8(3 * L) * 2 + 8;
    need = mark > need ? mark : need;
    need = coarsen > need ? coarsen : need;
    need = extract > need ? extract : need;
    return need;
}

int cfa_nvb_hash_build(const cfa_forest_t* f, int nv, cfa_stream_t stream) {
    if (!f || nv < 0 || f->hash_cap <= 0 || (f->hash_cap & (f->hash_cap - 1))) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_hash_clear, dim3(nblk(f->hash_cap)), dim3(BLOCK), 0, s, f->hash_cap, f->hkey, f->hval);
    if (nv > 0) hipLaunchKernelGGL(k_hash_fill, dim3(nblk(nv)), dim3(BLOCK), 0, s, *f, nv);
    return launch_status();
}

int cfa_nvb_leaves(const cfa_forest_t* f, int ne, int mode, const int* target, int* flags, int* list, void* tmp,
                   size_t tmp_bytes, cfa_stream_t stream) {
    if (!f || ne < 0 || mode < 0 || mode > 2 || (mode == 1 && !target) || !flags || !list || !tmp) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    if (ne == 0) {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(WAVE), 0, s, 1, f->cnt + CFA_NVB_LIST, 0);
        return launch_status();
    }
    hipLaunchKernelGGL(k_flag_leaves, dim3(nblk(ne)), dim3(BLOCK), 0, s, *f, ne, mode, target, flags);
    const hipError_t e = select_flagged(tmp, tmp_bytes, flags, list, f->cnt + CFA_NVB_LIST, ne, s);
    return e == hipSuccess ? launch_status() : (int)e;
}

int cfa_nvb_bisect(const cfa_forest_t* f, int n, const int* list, int inherit, int* target, double inlet_umax,
                   int* scratch, void* tmp, size_t tmp_bytes, cfa_stream_t stream) {
    if (!f || n < 0 || (n > 0 && (!list || !target || !scratch || !tmp))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    const hipStream_t s = as_stream(stream);
    int* mid = scratch;
    int* slot = mid + n;
    int* creator = slot + n;   // n + 1
    int* rank = creator + n + 1;   // n + 1: exclusive prefix of creator, rank[n] = midpoints created
    const dim3 g(nblk(n)), b(BLOCK);
    hipLaunchKernelGGL(k_bisect_claim, g, b, 0, s, *f, n, list, mid, slot);
    hipLaunchKernelGGL(k_bisect_creators, g, b, 0, s, *f, n, mid, slot, creator);
    const hipError_t e = scan_ints(tmp, tmp_bytes, creator, rank, n + 1, s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_bisect_vertices, g, b, 0, s, *f, n, list, slot, creator, rank, inlet_umax);
    hipLaunchKernelGGL(k_bisect_split, g, b, 0, s, *f, n, list, mid, slot, creator, inherit, target);
    hipLaunchKernelGGL(k_bisect_counts, dim3(1), dim3(WAVE), 0, s, *f, n, rank + n);
    return launch_status();
}

int cfa_nvb_mark(const cfa_forest_t* f, const cfa_nvb_mark_t* m, int ne, int nv, int* scratch, void* tmp,
                 size_t tmp_bytes, cfa_stream_t stream) {
    if (!f || !m || ne <= 0 || nv <= 0 || m->n_leaves <= 0 || !m->leaves || !m->eta || !m->av || !m->ai ||
        !m->target || !m->refine || !m->coarsen || !m->vflag || !m->nbr || !m->clist || !m->predicted || !scratch ||
        !tmp)
        return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    const int L = m->n_leaves;
    const dim3 gl(nblk(L)), b(BLOCK);
    // scratch layout
    int* w = scratch;
    auto take = [&](size_t ints) { int* p = w; w += align8(ints); return p; };
    double* src_v = reinterpret_cast<double*>(take(2 * (size_t)L));
    double* src_i = reinterpret_cast<double*>(take(2 * (size_t)L));
    int* inc_cnt = take((size_t)nv + 1);
    int* inc_ptr = take((size_t)nv + 1);
    int* inc_fill = take((size_t)nv);
    int* inc_list = take(3 * (size_t)L);
    int* flags = take(3 * (size_t)L);
    unsigned long long* keys_in = reinterpret_cast<unsigned long long*>(take(2 * (size_t)L));
    unsigned long long* keys_out = reinterpret_cast<unsigned long long*>(take(2 * (size_t)L));
    int* vals_in = take((size_t)L);
    int* vals_out = take((size_t)L);
    int* list = take((size_t)L);
    int* tgt = take((size_t)L);
    int* count = take(2);
    const cfa_nvb_mark_t mk = *m;
    // 1. normalised indicators on the forest slots
    GPU_TRY(hipMemsetAsync(m->av, 0, (size_t)ne * sizeof(double), s));
    GPU_TRY(hipMemsetAsync(m->ai, 0, (size_t)ne * sizeof(double), s));
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(ne)), b, 0, s, ne, m->target, -1);
    GPU_TRY(hipMemsetAsync(m->refine, 0, (size_t)ne, s));
    GPU_TRY(hipMemsetAsync(m->coarsen, 0, (size_t)ne, s));
    hipLaunchKernelGGL(k_normalise, gl, b, 0, s, *f, mk);
    // edge neighbours of the leaves
    GPU_TRY(hipMemsetAsync(inc_cnt, 0, ((size_t)nv + 1) * sizeof(int), s));
    hipLaunchKernelGGL(k_inc_count, gl, b, 0, s, *f, L, m->leaves, inc_cnt);
    GPU_TRY(scan_ints(tmp, tmp_bytes, inc_cnt, inc_ptr, nv + 1, s));
    GPU_TRY(hipMemcpyAsync(inc_fill, inc_ptr, (size_t)nv * sizeof(int), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_inc_fill, gl, b, 0, s, *f, L, m->leaves, inc_fill, inc_list);
    hipLaunchKernelGGL(k_leaf_nbr, gl, b, 0, s, *f, L, m->leaves, inc_ptr, inc_list, m->nbr);
    // 2. look-ahead
    if (m->lookahead_time > 0.0) {
        hipLaunchKernelGGL(k_sources, gl, b, 0, s, mk, src_v, src_i);
        hipLaunchKernelGGL(k_lookahead, dim3((L + 63) / 64), dim3(64), 0, s, *f, mk, src_v, src_i, m->nbr);
    }
    // 3. refinement marks, largest score first (stable: ties in leaf order)
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(WAVE), 0, s, 1, count, 0);
    hipLaunchKernelGGL(k_score_keys, gl, b, 0, s, mk, keys_in, vals_in, count);
    {
        size_t tb = tmp_bytes;
        GPU_TRY(rocprim::radix_sort_pairs(tmp, tb, keys_in, keys_out, vals_in, vals_out, (size_t)L, 0, 64, s));
    }
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(WAVE), 0, s, count, f->cnt + CFA_NVB_ORDERED);
    hipLaunchKernelGGL(k_order_targets, gl, b, 0, s, *f, mk, vals_out, count, list, tgt);
    hipLaunchKernelGGL(k_budget, dim3(1), dim3(256), 0, s, *f, mk, count, list, tgt, 0);
    // minimum level
    if (m->min_level > 0) {
        hipLaunchKernelGGL(k_candidates, gl, b, 0, s, *f, mk, 0, m->nbr, flags);
        GPU_TRY(select_flagged(tmp, tmp_bytes, flags, list, count, L, s));
        hipLaunchKernelGGL(k_leaf_slots, gl, b, 0, s, mk, count, list);
        hipLaunchKernelGGL(k_budget, dim3(1), dim3(256), 0, s, *f, mk, count, list, nullptr, 2 * m->min_level);
    }
    // 4. hole filling, two passes
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_candidates, gl, b, 0, s, *f, mk, 1, m->nbr, flags);
        GPU_TRY(select_flagged(tmp, tmp_bytes, flags, list, count, L, s));
        hipLaunchKernelGGL(k_leaf_slots, gl, b, 0, s, mk, count, list);
        hipLaunchKernelGGL(k_plus_two, gl, b, 0, s, *f, count, list, tgt);
        hipLaunchKernelGGL(k_budget, dim3(1), dim3(256), 0, s, *f, mk, count, list, tgt, 0);
    }
    // 5. buffer layers, then the protected layer
    for (int layer = 0; layer < m->buffer_layers; ++layer) {
        GPU_TRY(hipMemsetAsync(m->vflag, 0, (size_t)nv, s));
        hipLaunchKernelGGL(k_vflag_marked, gl, b, 0, s, *f, mk);
        hipLaunchKernelGGL(k_candidates, gl, b, 0, s, *f, mk, 2, m->nbr, flags);
        GPU_TRY(select_flagged(tmp, tmp_bytes, flags, list, count, L, s));
        hipLaunchKernelGGL(k_leaf_slots, gl, b, 0, s, mk, count, list);
        hipLaunchKernelGGL(k_plus_two, gl, b, 0, s, *f, count, list, tgt);
        hipLaunchKernelGGL(k_budget, dim3(1), dim3(256), 0, s, *f, mk, count, list, tgt, 0);
    }
    GPU_TRY(hipMemsetAsync(m->vflag, 0, (size_t)nv, s));
    hipLaunchKernelGGL(k_vflag_marked, gl, b, 0, s, *f, mk);
    // 6. coarsening flags, the list in leaf order
    hipLaunchKernelGGL(k_coarsen_flags, gl, b, 0, s, *f, mk, m->nbr, flags);
    GPU_TRY(select_flagged(tmp, tmp_bytes, flags, m->clist, f->cnt + CFA_NVB_LIST, L, s));
    hipLaunchKernelGGL(k_leaf_slots, gl, b, 0, s, mk, f->cnt + CFA_NVB_LIST, m->clist);
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(WAVE), 0, s, f->cnt + CFA_NVB_LIST, f->cnt + CFA_NVB_MARKED_COARSEN);
    return launch_status();
}

int cfa_nvb_coarsen_sweep(const cfa_forest_t* f, const cfa_nvb_mark_t* m, int nv, int* scratch, void* tmp,
                          size_t tmp_bytes, cfa_stream_t stream) {
    if (!f || !m || nv <= 0 || m->n_leaves <= 0 || !m->clist || !scratch || !tmp) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    int* w = scratch;
    auto take = [&](size_t ints) { int* p = w; w += align8(ints); return p; };
    int* vcand = take((size_t)nv + 1);
    int* go = take((size_t)nv + 1);
    int* sizes = take((size_t)nv + 1);
    int* off = take((size_t)nv + 1);
    int* count = take(2);
    const cfa_nvb_mark_t mk = *m;
    GPU_TRY(hipMemsetAsync(vcand, 0, (size_t)nv * sizeof(int), s));
    // the list holds at most L leaves (the first sweep) or 2 merged parents per removed vertex
    const int n_max = m->n_leaves > 2 * nv ? m->n_leaves : 2 * nv;
    hipLaunchKernelGGL(k_collapse_check, dim3(nblk(n_max)), dim3(BLOCK), 0, s, *f, mk, n_max, f->cnt + CFA_NVB_LIST,
                       vcand);
    GPU_TRY(select_flagged(tmp, tmp_bytes, vcand, go, count, nv, s));
    hipLaunchKernelGGL(k_collapse_sizes, dim3(nblk(nv + 1)), dim3(BLOCK), 0, s, *f, go, count, nv + 1, sizes);
    GPU_TRY(scan_ints(tmp, tmp_bytes, sizes, off, nv + 1, s));
    hipLaunchKernelGGL(k_collapse, dim3(nblk(nv)), dim3(BLOCK), 0, s, *f, mk, go, count, off);
    hipLaunchKernelGGL(k_collapse_counts, dim3(1), dim3(WAVE), 0, s, *f, count, off);
    return launch_status();
}

int cfa_nvb_extract(const cfa_forest_t* f, int nv, int n_leaves, const int* leaves, int* scratch, void* tmp,
                    size_t tmp_bytes, int* out_leaf, int* out_node, int* elem_v, int* gen, double* x, double* y,
                    int* tag, double* values, cfa_stream_t stream) {
    if (!f || nv <= 0 || n_leaves <= 0 || !leaves || !scratch || !tmp || !out_leaf || !out_node || !elem_v || !gen ||
        !x || !y || !tag || !values)
        return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    const int L = n_leaves;
    int* w = scratch;
    auto take = [&](size_t ints) { int* p = w; w += align8(ints); return p; };
    unsigned* keys = reinterpret_cast<unsigned*>(take((size_t)L));
    unsigned* keys_sorted = reinterpret_cast<unsigned*>(take((size_t)L));
    int* first = take((size_t)nv);
    int* compact = take((size_t)nv);
    int* flags = take(3 * (size_t)L);
    int* pos = take(3 * (size_t)L);
    const dim3 b(BLOCK);
    hipLaunchKernelGGL(k_morton, dim3(nblk(L)), b, 0, s, *f, L, leaves, keys);
    {
        size_t tb = tmp_bytes;
        GPU_TRY(rocprim::radix_sort_pairs(tmp, tb, keys, keys_sorted, leaves, out_leaf, (size_t)L, 0, 32, s));
    }
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nv)), b, 0, s, nv, first, H_UNSET);
    hipLaunchKernelGGL(k_first_use, dim3(nblk(3 * (long long)L)), b, 0, s, *f, L, out_leaf, first);
    hipLaunchKernelGGL(k_first_flags, dim3(nblk(3 * (long long)L)), b, 0, s, *f, L, out_leaf, first, flags);
    GPU_TRY(select_flagged(tmp, tmp_bytes, flags, pos, f->cnt + CFA_NVB_NODES, 3 * L, s));
    hipLaunchKernelGGL(k_nodes, dim3(nblk(nv)), b, 0, s, *f, f->cnt + CFA_NVB_NODES, nv, out_leaf, pos, out_node,
                       compact, x, y, tag, values);
    hipLaunchKernelGGL(k_elements, dim3(nblk(L)), b, 0, s, *f, L, out_leaf, compact, elem_v, gen);
    return launch_status();
}

int cfa_nvb_uniform_targets(const cfa_forest_t* f, int ne, int n_leaves, const int* leaves, int levels, int* target,
                            cfa_stream_t stream) {
    if (!f || ne <= 0 || n_leaves < 0 || (n_leaves > 0 && !leaves) || !target) return CFA_E_ARG;
    const hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(ne)), dim3(BLOCK), 0, s, ne, target, -1);
    if (n_leaves > 0)
        hipLaunchKernelGGL(k_uniform_targets, dim3(nblk(n_leaves)), dim3(BLOCK), 0, s, *f, n_leaves, leaves, levels,
                           target);
    return launch_status();
}

int cfa_nvb_scatter(const cfa_forest_t* f, int n, const int* nodes, const double* values, cfa_stream_t stream) {
    if (!f || n < 0 || (n > 0 && (!nodes || !values))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(BLOCK), 0, as_stream(stream), *f, n, nodes, values);
    return launch_status();
}

int cfa_nvb_set_xy(const cfa_forest_t* f, int n, const int* slots, const double* xy, cfa_stream_t stream) {
    if (!f || n < 0 || (n > 0 && (!slots || !xy))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_set_xy, dim3(nblk(n)), dim3(BLOCK), 0, as_stream(stream), *f, n, slots, xy);
    return launch_status();
}

int cfa_nvb_replay(const cfa_forest_t* f, int first, int last, double inlet_umax, cfa_stream_t stream) {
    if (!f || first < 0 || last < first) return CFA_E_ARG;
    if (last == first) return CFA_OK;
    hipLaunchKernelGGL(k_replay, dim3(nblk(last - first)), dim3(BLOCK), 0, as_stream(stream), *f, first, last,
                       inlet_umax);
    return launch_status();
}

int cfa_nvb_gather(const cfa_forest_t* f, int n, const int* nodes, double* values, cfa_stream_t stream) {
    if (!f || n < 0 || (n > 0 && (!nodes || !values))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_gather, dim3(nblk(n)), dim3(BLOCK), 0, as_stream(stream), *f, n, nodes, values);
    return launch_status();
}

}  // extern "C"
