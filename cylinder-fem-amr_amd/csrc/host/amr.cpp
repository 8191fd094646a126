// Host side of the device NVB forest: owns the forest's device arrays (growing with the same
// x1.5 rule as the solver's buffers), sequences the kernel groups of cfa_nvb_* for one adaptation
// (marking, two coarsening sweeps, refinement rounds each closed by hanging-edge passes), the
// extraction and the transfers, and reads the device counters back once per refinement pass.
// Reference interface: src/cfa/amr.cuh:58-471 (AdaptiveMesh).
#include "amr.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gpu.hpp"

namespace cfa {

namespace {

// Device array on a given stream whose growth keeps the first `used` elements.
template <typename T>
struct Buf {
    T* ptr = nullptr;
    size_t cap = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (ptr) cfa_free(ptr); }
    void ensure(size_t need, size_t used, cfa_stream_t s) {
        if (need <= cap && ptr) return;
        const size_t c = need + need / 2 + 256;
        void* p = nullptr;
        GPU_CHECK(cfa_malloc(&p, c * sizeof(T)));
        if (ptr && used) GPU_CHECK(cfa_memcpy(p, ptr, std::min(used, cap) * sizeof(T), 2, s));
        if (ptr) {
            GPU_CHECK(cfa_stream_sync(s));
            cfa_free(ptr);
        }
        ptr = static_cast<T*>(p);
        cap = c;
    }
    void put(const T* h, size_t n, cfa_stream_t s) {
        ensure(n, 0, s);
        if (n) GPU_CHECK(cfa_memcpy(ptr, h, n * sizeof(T), 0, s));
    }
    void get(T* h, size_t n, cfa_stream_t s) const {
        if (n) GPU_CHECK(cfa_memcpy(h, ptr, n * sizeof(T), 1, s));
    }
};

size_t pow2_at_least(size_t n) {
    size_t p = 1024;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

struct AdaptiveMesh::Device {
    cfa_stream_t s = nullptr;
    // forest
    Buf<int> ev, egen, eparent, ech, efree;
    Buf<unsigned char> ealive;
    Buf<double> vx, vy, vu, vv, vp, vuo, vvo;
    Buf<int> vtag, vpar, vfree, log, cyl_log, cyl_slots;
    Buf<double> cyl_xy;
    Buf<unsigned char> valive;
    Buf<unsigned long long> vedge, hkey;
    Buf<int> hval, cnt;
    size_t hash_cap = 0;
    // marking
    Buf<double> av, ai, eta;
    Buf<int> target, nbr, clist;
    Buf<unsigned char> refine, coarsen, vflag;
    Buf<long long> predicted;
    // lists, extraction, scratch
    Buf<int> leaves, nodes, flags, list, scratch, elem_v, gen, tag, old_nodes;
    Buf<double> x, y, values;
    Buf<char> tmp;
    size_t tmp_bytes = 0;
    // host mirrors of the counters
    int c[CFA_NVB_COUNTERS] = {};
    int n_leaves = 0, n_nodes = 0;
    std::vector<int> passes;   // log offsets of the bisection passes of the last adaptation
    std::vector<int> pass_n;   // bisections per pass of the last adaptation
    // host copy of the cylinder vertices' coordinates (by vertex slot): the arc positions of new
    // cylinder midpoints are computed here with the host's atan2 / cos / sin (amr.cuh:802-846)
    std::vector<double> cx, cy;
    int cyl_done = 0;   // cyl_log entries already placed

    cfa_forest_t view() const {
        cfa_forest_t f;
        f.ev = ev.ptr;  f.egen = egen.ptr;  f.eparent = eparent.ptr;  f.ech = ech.ptr;  f.ealive = ealive.ptr;
        f.vx = vx.ptr;  f.vy = vy.ptr;  f.vu = vu.ptr;  f.vv = vv.ptr;  f.vp = vp.ptr;  f.vuo = vuo.ptr;
        f.vvo = vvo.ptr;  f.vtag = vtag.ptr;  f.valive = valive.ptr;  f.vpar = vpar.ptr;  f.vedge = vedge.ptr;
        f.efree = efree.ptr;  f.vfree = vfree.ptr;  f.log = log.ptr;  f.cyl_log = cyl_log.ptr;
        f.hkey = hkey.ptr;  f.hval = hval.ptr;  f.hash_cap = (int)hash_cap;
        f.cnt = cnt.ptr;
        return f;
    }
    void read_counters() { cnt.get(c, CFA_NVB_COUNTERS, s); }
    void write_counters() { GPU_CHECK(cfa_memcpy(cnt.ptr, c, sizeof(c), 0, s)); }

    // room for `ne` element slots and `nv` vertex slots, keeping the used ones
    void reserve_elements(size_t ne) {
        const size_t used = (size_t)c[CFA_NVB_NE];
        ev.ensure(3 * ne, 3 * used, s);
        egen.ensure(ne, used, s);
        eparent.ensure(ne, used, s);
        ech.ensure(2 * ne, 2 * used, s);
        ealive.ensure(ne, used, s);
        efree.ensure(ne, (size_t)c[CFA_NVB_EFREE], s);
        target.ensure(ne, used, s);
        flags.ensure(ne, 0, s);
        list.ensure(ne, (size_t)c[CFA_NVB_LIST], s);   // keeps the current selection
    }
    void reserve_vertices(size_t nv) {
        const size_t used = (size_t)c[CFA_NVB_NV];
        for (Buf<double>* b : {&vx, &vy, &vu, &vv, &vp, &vuo, &vvo}) b->ensure(nv, used, s);
        vtag.ensure(nv, used, s);
        valive.ensure(nv, used, s);
        vpar.ensure(2 * nv, 2 * used, s);
        vedge.ensure(nv, used, s);
        vfree.ensure(nv, (size_t)c[CFA_NVB_VFREE], s);
    }
    void reserve_scratch(size_t ne, size_t nv, size_t L) {
        scratch.ensure(cfa_nvb_scratch_ints((int)ne, (int)nv, (int)L), 0, s);
        const size_t need = cfa_nvb_tmp_bytes((int)std::max({ne, nv + 1, 3 * L}));
        if (need > tmp_bytes) {
            tmp.ensure(need, 0, s);
            tmp_bytes = tmp.cap;
        }
    }
    // the table of bisected edges, rebuilt with room for `midpoints` entries at half load
    void rebuild_hash(size_t midpoints) {
        const size_t cap = pow2_at_least(4 * midpoints);
        if (cap > hash_cap) {
            hkey.ensure(cap, 0, s);
            hval.ensure(cap, 0, s);
            hash_cap = cap;
        }
        const cfa_forest_t f = view();
        GPU_CHECK(cfa_nvb_hash_build(&f, c[CFA_NVB_NV], s));
    }
    // list of element slots by mode (0 leaves, 1 below target, 2 hanging); returns its length
    int select(int mode) {
        const cfa_forest_t f = view();
        GPU_CHECK(cfa_nvb_leaves(&f, c[CFA_NVB_NE], mode, target.ptr, flags.ptr, list.ptr, tmp.ptr, tmp_bytes, s));
        read_counters();
        if (c[CFA_NVB_ERROR])
            fatal("[AMR] internal error: an edge bisected by more than two elements, a full edge table or a "
                  "bisection list entry that is no element slot");
        place_cylinder_midpoints();
        return c[CFA_NVB_LIST];
    }
    // arc positions of the cylinder midpoints the last pass created (before any pass reads them)
    void place_cylinder_midpoints() {
        const int total = c[CFA_NVB_CYLINDER], k = total - cyl_done;
        if (k <= 0) return;
        std::vector<int> e(3 * (size_t)k), slots(k);
        GPU_CHECK(cfa_memcpy(e.data(), cyl_log.ptr + 3 * (size_t)cyl_done, e.size() * sizeof(int), 1, s));
        std::vector<double> xy(2 * (size_t)k);
        for (int q = 0; q < k; ++q) {
            const int m = e[3 * q], a = e[3 * q + 1], b = e[3 * q + 2];
            const double mx = 0.5 * (cx[a] + cx[b]), my = 0.5 * (cy[a] + cy[b]);
            const double th = std::atan2(my - Sim::Y_CYL, mx - Sim::X_CYL);
            if ((size_t)m >= cx.size()) {
                cx.resize(m + 1 + m / 2);
                cy.resize(cx.size());
            }
            cx[m] = Sim::X_CYL + Sim::R_CYL * std::cos(th);
            cy[m] = Sim::Y_CYL + Sim::R_CYL * std::sin(th);
            slots[q] = m;
            xy[2 * q] = cx[m];
            xy[2 * q + 1] = cy[m];
        }
        cyl_slots.put(slots.data(), k, s);
        cyl_xy.put(xy.data(), xy.size(), s);
        const cfa_forest_t f = view();
        GPU_CHECK(cfa_nvb_set_xy(&f, k, cyl_slots.ptr, cyl_xy.ptr, s));
        cyl_done = total;
    }
    // one pass of n bisections of `list`; `occupied` counts the table's used entries
    void bisect(int n, bool inherit, double umax, size_t& occupied) {
        const size_t ne = (size_t)c[CFA_NVB_NE] + 2 * (size_t)n, nv = (size_t)c[CFA_NVB_NV] + n;
        reserve_elements(ne);
        reserve_vertices(nv);
        log.ensure(3 * ((size_t)c[CFA_NVB_LOG] + n), 3 * (size_t)c[CFA_NVB_LOG], s);
        cyl_log.ensure(3 * ((size_t)c[CFA_NVB_CYLINDER] + n), 3 * (size_t)c[CFA_NVB_CYLINDER], s);
        reserve_scratch(ne, nv, (size_t)n_leaves);
        if (occupied + n > hash_cap / 2) {   // rebuilt with room for every slot (drops removed edges)
            rebuild_hash(nv);
            occupied = (size_t)c[CFA_NVB_NV];
        }
        occupied += n;
        const cfa_forest_t f = view();
        GPU_CHECK(cfa_nvb_bisect(&f, n, list.ptr, inherit ? 1 : 0, target.ptr, umax, scratch.ptr, tmp.ptr, tmp_bytes,
                                 s));
        // element slots in use after the pass, as the device counts them (2n children, recycled
        // slots first), so that the next selection scans the new children too
        const int etake = std::min(2 * n, c[CFA_NVB_EFREE]);
        c[CFA_NVB_NE] += 2 * n - etake;
        c[CFA_NVB_EFREE] -= etake;
    }
    // Refinement rounds (amr.cuh:476-508): every leaf below its target generation is bisected
    // (children inherit the target), then every hanging leaf, pass after pass, until none hangs.
    // passes: the log offset at the end of each pass.
    void refine_rounds(double umax, AmrStats& st) {
        size_t occupied = (size_t)c[CFA_NVB_NV];
        passes.assign(1, c[CFA_NVB_LOG]);
        pass_n.clear();
        while (true) {
            const int n = select(1);
            if (n == 0) break;
            pass_n.push_back(n);
            bisect(n, true, umax, occupied);
            st.bisections += n;
            while (true) {
                const int h = select(2);   // also reads the counters of the pass before
                passes.push_back(c[CFA_NVB_LOG]);
                if (h == 0) break;
                pass_n.push_back(-h);
                bisect(h, false, umax, occupied);
                st.closure_bisections += h;
            }
        }
        if (passes.back() != c[CFA_NVB_LOG]) passes.push_back(c[CFA_NVB_LOG]);
    }
};

AdaptiveMesh::AdaptiveMesh(const Mesh2D& base, const std::vector<double>& u0, const std::vector<double>& v0,
                           const std::vector<double>& p0)
    : d_(new Device) {
    GPU_CHECK(cfa_init(0));
    GPU_CHECK(cfa_stream_create(&d_->s));
    Device& d = *d_;
    const int ne = base.num_elems, nv = base.num_nodes;
    d.cnt.ensure(CFA_NVB_COUNTERS, 0, d.s);
    d.reserve_elements(ne);
    d.reserve_vertices(nv);
    std::vector<int> ev(3 * (size_t)ne), zero_i(ne, 0), neg(2 * (size_t)std::max(ne, nv), -1);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) ev[3 * e + k] = base.elems[e].v[k];
    d.ev.put(ev.data(), ev.size(), d.s);
    d.egen.put(zero_i.data(), ne, d.s);
    d.eparent.put(neg.data(), ne, d.s);
    d.ech.put(neg.data(), 2 * (size_t)ne, d.s);
    const std::vector<unsigned char> ones(std::max(ne, nv), 1);
    d.ealive.put(ones.data(), ne, d.s);
    d.vx.put(base.x.data(), nv, d.s);
    d.vy.put(base.y.data(), nv, d.s);
    d.vu.put(u0.data(), nv, d.s);
    d.vv.put(v0.data(), nv, d.s);
    d.vp.put(p0.data(), nv, d.s);
    d.vuo.put(u0.data(), nv, d.s);
    d.vvo.put(v0.data(), nv, d.s);
    d.vtag.put(base.bc_tag.data(), nv, d.s);
    d.valive.put(ones.data(), nv, d.s);
    d.vpar.put(neg.data(), 2 * (size_t)nv, d.s);
    const std::vector<unsigned long long> zero_k(nv, 0ull);
    d.vedge.put(zero_k.data(), nv, d.s);
    d.c[CFA_NVB_NE] = ne;
    d.c[CFA_NVB_NV] = nv;
    d.write_counters();
    d.cx.assign(nv, 0.0);
    d.cy.assign(nv, 0.0);
    for (int i = 0; i < nv; ++i)
        if (base.bc_tag[i] == BC_CYLINDER) {
            d.cx[i] = base.x[i];
            d.cy[i] = base.y[i];
        }
    GPU_CHECK(cfa_stream_sync(d.s));
}

AdaptiveMesh::~AdaptiveMesh() {
    if (d_ && d_->s) {
        cfa_stream_sync(d_->s);
        cfa_stream_t s = d_->s;
        d_.reset();   // frees the arrays (cfa_free synchronises the device)
        cfa_stream_destroy(s);
    }
}

int AdaptiveMesh::num_leaves() const { return d_->n_leaves; }
int AdaptiveMesh::num_vertex_slots() const { return d_->c[CFA_NVB_NV]; }

const std::vector<int>& AdaptiveMesh::pass_sizes() const { return d_->pass_n; }

void AdaptiveMesh::slot_counts(int out[4]) const {
    out[0] = d_->c[CFA_NVB_NE];
    out[1] = d_->c[CFA_NVB_NV];
    out[2] = d_->c[CFA_NVB_EFREE];
    out[3] = d_->c[CFA_NVB_VFREE];
}

void AdaptiveMesh::extract(Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                           std::vector<double>& h_p, std::vector<double>& h_uo, std::vector<double>& h_vo) {
    Device& d = *d_;
    auto t = Clock::now();
    const size_t ne = (size_t)d.c[CFA_NVB_NE], nv = (size_t)d.c[CFA_NVB_NV];
    d.reserve_scratch(ne, nv, ne);
    const int L = d.select(0);
    d.n_leaves = L;
    dbg_ms[D_LEAVES] += ms_since(t);
    t = Clock::now();
    d.leaves.ensure(L, 0, d.s);
    d.nodes.ensure(nv, 0, d.s);
    d.elem_v.ensure(3 * (size_t)L, 0, d.s);
    d.gen.ensure(L, 0, d.s);
    d.x.ensure(nv, 0, d.s);
    d.y.ensure(nv, 0, d.s);
    d.tag.ensure(nv, 0, d.s);
    d.values.ensure(5 * nv, 0, d.s);
    d.reserve_scratch(ne, nv, L);
    const cfa_forest_t f = d.view();
    GPU_CHECK(cfa_nvb_extract(&f, (int)nv, L, d.list.ptr, d.scratch.ptr, d.tmp.ptr, d.tmp_bytes, d.leaves.ptr,
                              d.nodes.ptr, d.elem_v.ptr, d.gen.ptr, d.x.ptr, d.y.ptr, d.tag.ptr, d.values.ptr, d.s));
    d.read_counters();
    const int nn = d.c[CFA_NVB_NODES];
    d.n_nodes = nn;
    dbg_ms[D_EXTRACT] += ms_since(t);
    t = Clock::now();
    mesh.num_nodes = nn;
    mesh.num_elems = L;
    mesh.x.resize(nn);
    mesh.y.resize(nn);
    mesh.bc_tag.resize(nn);
    d.x.get(mesh.x.data(), nn, d.s);
    d.y.get(mesh.y.data(), nn, d.s);
    d.tag.get(mesh.bc_tag.data(), nn, d.s);
    std::vector<int> ev(3 * (size_t)L), gen(L);
    d.elem_v.get(ev.data(), ev.size(), d.s);
    d.gen.get(gen.data(), L, d.s);
    mesh.elems.resize(L);
    for (int i = 0; i < L; ++i) {
        for (int k = 0; k < 3; ++k) mesh.elems[i].v[k] = ev[3 * (size_t)i + k];
        mesh.elems[i].gen = gen[i];
        mesh.elems[i].level = gen[i] / 2;
    }
    std::vector<double>* out[5] = {&h_u, &h_v, &h_p, &h_uo, &h_vo};
    for (int k = 0; k < 5; ++k) {
        out[k]->resize(nn);
        GPU_CHECK(cfa_memcpy(out[k]->data(), d.values.ptr + (size_t)k * nn, nn * sizeof(double), 1, d.s));
    }
    vert_ids_.resize(nn);
    d.nodes.get(vert_ids_.data(), nn, d.s);
    leaf_ids_.resize(L);
    d.leaves.get(leaf_ids_.data(), L, d.s);
    dbg_ms[D_DOWNLOAD] += ms_since(t);
}

void AdaptiveMesh::scatter_solution(const std::vector<double>& h_u, const std::vector<double>& h_v,
                                    const std::vector<double>& h_p, const std::vector<double>& h_uo,
                                    const std::vector<double>& h_vo) {
    Device& d = *d_;
    const size_t n = vert_ids_.size();
    d.values.ensure(5 * n, 0, d.s);
    const std::vector<double>* in[5] = {&h_u, &h_v, &h_p, &h_uo, &h_vo};
    for (int k = 0; k < 5; ++k) GPU_CHECK(cfa_memcpy(d.values.ptr + k * n, in[k]->data(), n * sizeof(double), 0, d.s));
    const cfa_forest_t f = d.view();
    GPU_CHECK(cfa_nvb_scatter(&f, (int)n, d.nodes.ptr, d.values.ptr, d.s));
    GPU_CHECK(cfa_stream_sync(d.s));
}

void AdaptiveMesh::retransfer(const std::vector<int>& old_vids, const std::vector<double>& h_u,
                              const std::vector<double>& h_v, const std::vector<double>& h_p,
                              const std::vector<double>& h_uo, const std::vector<double>& h_vo,
                              std::vector<double>& n_u, std::vector<double>& n_v, std::vector<double>& n_p,
                              std::vector<double>& n_uo, std::vector<double>& n_vo) {
    Device& d = *d_;
    const size_t n_old = old_vids.size(), nn = vert_ids_.size();
    d.old_nodes.put(old_vids.data(), n_old, d.s);
    d.values.ensure(5 * std::max(n_old, nn), 0, d.s);
    const std::vector<double>* in[5] = {&h_u, &h_v, &h_p, &h_uo, &h_vo};
    for (int k = 0; k < 5; ++k)
        GPU_CHECK(cfa_memcpy(d.values.ptr + k * n_old, in[k]->data(), n_old * sizeof(double), 0, d.s));
    const cfa_forest_t f = d.view();
    GPU_CHECK(cfa_nvb_scatter(&f, (int)n_old, d.old_nodes.ptr, d.values.ptr, d.s));
    // the midpoints of the last adaptation, pass after pass (a pass only reads earlier values)
    for (size_t q = 0; q + 1 < d.passes.size(); ++q)
        GPU_CHECK(cfa_nvb_replay(&f, d.passes[q], d.passes[q + 1], inlet_umax, d.s));
    GPU_CHECK(cfa_nvb_gather(&f, (int)nn, d.nodes.ptr, d.values.ptr, d.s));
    std::vector<double>* out[5] = {&n_u, &n_v, &n_p, &n_uo, &n_vo};
    for (int k = 0; k < 5; ++k) {
        out[k]->resize(nn);
        GPU_CHECK(cfa_memcpy(out[k]->data(), d.values.ptr + k * nn, nn * sizeof(double), 1, d.s));
    }
}


AmrStats AdaptiveMesh::refine_uniform(int levels) {
    Device& d = *d_;
    AmrStats st;
    st.leaves_before = d.n_leaves;
    d.read_counters();
    d.c[CFA_NVB_LOG] = d.c[CFA_NVB_CYLINDER] = 0;
    d.cyl_done = 0;
    d.write_counters();
    const size_t ne = (size_t)d.c[CFA_NVB_NE], nv = (size_t)d.c[CFA_NVB_NV];
    d.reserve_elements(ne);
    d.reserve_scratch(ne, nv, std::max(d.n_leaves, 1));
    d.rebuild_hash(nv);
    const cfa_forest_t f = d.view();
    GPU_CHECK(cfa_nvb_uniform_targets(&f, (int)ne, d.n_leaves, d.leaves.ptr, levels, d.target.ptr, d.s));
    auto t = Clock::now();
    d.refine_rounds(inlet_umax, st);
    dbg_ms[D_REFINE] += ms_since(t);
    st.leaves_after = d.select(0);
    return st;
}

AmrStats AdaptiveMesh::adapt(const std::vector<double>& eta, const AmrParams& prm) {
    Device& d = *d_;
    const auto t_start = Clock::now();
    AmrStats st;
    const int L = d.n_leaves;
    st.leaves_before = L;
    // the log and the adaptation's counters start from zero
    d.read_counters();
    const size_t ne = (size_t)d.c[CFA_NVB_NE], nv = (size_t)d.c[CFA_NVB_NV];
    d.c[CFA_NVB_LOG] = d.c[CFA_NVB_CYLINDER] = 0;
    d.cyl_done = 0;
    d.c[CFA_NVB_MARKED_REFINE] = d.c[CFA_NVB_MARKED_COARSEN] = d.c[CFA_NVB_REMOVED] = 0;
    d.write_counters();
    d.reserve_elements(ne);
    d.reserve_scratch(ne, nv, L);
    d.rebuild_hash(nv);
    d.eta.put(eta.data(), 2 * (size_t)L, d.s);
    d.av.ensure(ne, 0, d.s);
    d.ai.ensure(ne, 0, d.s);
    d.refine.ensure(ne, 0, d.s);
    d.coarsen.ensure(ne, 0, d.s);
    d.vflag.ensure(nv, 0, d.s);
    d.nbr.ensure(3 * (size_t)L, 0, d.s);
    d.clist.ensure(std::max((size_t)L, 2 * nv), 0, d.s);
    const long long predicted = L;
    d.predicted.put(&predicted, 1, d.s);

    cfa_nvb_mark_t m;
    m.n_leaves = L;
    m.leaves = d.leaves.ptr;
    m.eta = d.eta.ptr;
    m.max_level = prm.max_level;
    m.min_level = prm.min_level;
    m.buffer_layers = prm.buffer_layers;
    m.max_elems = prm.max_elems;
    m.tol_vort = prm.tol_vort;
    m.tol_interp = prm.tol_interp;
    m.coarsen_safety = prm.coarsen_safety;
    m.lookahead_time = prm.lookahead_time;
    m.keep_radius = prm.keep_radius;
    m.tol_radius = prm.tol_radius;
    m.tol_power = prm.tol_power;
    m.av = d.av.ptr;
    m.ai = d.ai.ptr;
    m.target = d.target.ptr;
    m.refine = d.refine.ptr;
    m.coarsen = d.coarsen.ptr;
    m.vflag = d.vflag.ptr;
    m.nbr = d.nbr.ptr;
    m.clist = d.clist.ptr;
    m.predicted = d.predicted.ptr;
    const cfa_forest_t f = d.view();
    auto t = Clock::now();
    GPU_CHECK(cfa_nvb_mark(&f, &m, (int)ne, (int)nv, d.scratch.ptr, d.tmp.ptr, d.tmp_bytes, d.s));
    GPU_CHECK(cfa_stream_sync(d.s));
    st.ms_mark = ms_since(t);
    dbg_ms[D_MARK] += st.ms_mark;
    t = Clock::now();
    for (int sweep = 0; sweep < 2; ++sweep)
        GPU_CHECK(cfa_nvb_coarsen_sweep(&f, &m, (int)nv, d.scratch.ptr, d.tmp.ptr, d.tmp_bytes, d.s));
    d.read_counters();
    st.ms_coarsen = ms_since(t);
    dbg_ms[D_COARSEN] += st.ms_coarsen;
    t = Clock::now();
    d.refine_rounds(inlet_umax, st);
    st.ms_refine = ms_since(t);
    dbg_ms[D_REFINE] += st.ms_refine;
    st.marked_refine = d.c[CFA_NVB_MARKED_REFINE];
    st.marked_coarsen = d.c[CFA_NVB_MARKED_COARSEN];
    st.removed_vertices = d.c[CFA_NVB_REMOVED];
    st.leaves_after = d.select(0);
    st.elapsed_ms = ms_since(t_start);
    return st;
}

}  // namespace cfa
