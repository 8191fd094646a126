// ORACLE (test infrastructure only): serial restatement of the adaptive-mesh rules of
// src/cfa/amr.cuh. Plain loops over the whole forest (every closure pass re-examines every leaf,
// every work list is a scan in slot order), an ordered map for the bisected edges, glibc libm for
// the geometry, as the reference's host code uses it.
#include "cfa_oracle_amr.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace oracle {

using cfa::Sim;

namespace {

// src/cfa/amr.cuh:566-581: 16 bits per axis, interleaved (x in the even bits)
uint32_t interleave16(uint32_t n) {
    uint32_t r = 0;
    for (int b = 0; b < 16; ++b) r |= ((n >> b) & 1u) << (2 * b);
    return r;
}
uint32_t morton(double sx, double sy) {
    auto q = [](double s) { return (uint32_t)std::min(65535.0, std::max(0.0, s * 65535.0)); };
    return interleave16(q(sx)) | (interleave16(q(sy)) << 1);
}

}  // namespace

NvbForest::NvbForest(const cfa::Mesh2D& base, const std::vector<double>& u0, const std::vector<double>& v0,
                     const std::vector<double>& p0)
    : x_(base.x), y_(base.y), u_(u0), v_(v0), p_(p0), uo_(u0), vo_(v0), tag_(base.bc_tag) {
    const int nv = base.num_nodes;
    valive_.assign(nv, 1);
    vpar_.assign(nv, {-1, -1});
    vedge_.assign(nv, 0);
    el_.resize(base.num_elems);
    for (int e = 0; e < base.num_elems; ++e) {
        for (int k = 0; k < 3; ++k) el_[e].v[k] = base.elems[e].v[k];
        el_[e].alive = true;
    }
}

double NvbForest::area(int e) const {
    const int* w = el_[e].v;
    return 0.5 * std::fabs((x_[w[1]] - x_[w[0]]) * (y_[w[2]] - y_[w[0]]) -
                           (x_[w[2]] - x_[w[0]]) * (y_[w[1]] - y_[w[0]]));
}

// src/cfa/amr.cuh:625-634
bool NvbForest::inside(int e, double px, double py, double l[3]) const {
    const int* w = el_[e].v;
    const double x0 = x_[w[0]], y0 = y_[w[0]];
    const double d = (x_[w[1]] - x0) * (y_[w[2]] - y0) - (x_[w[2]] - x0) * (y_[w[1]] - y0);
    l[1] = ((px - x0) * (y_[w[2]] - y0) - (x_[w[2]] - x0) * (py - y0)) / d;
    l[2] = ((x_[w[1]] - x0) * (py - y0) - (px - x0) * (y_[w[1]] - y0)) / d;
    l[0] = 1.0 - l[1] - l[2];
    return l[1] >= -1.0e-12 && l[2] >= -1.0e-12 && l[1] + l[2] <= 1.0 + 1.0e-12;
}

// ---- extraction (src/cfa/amr.cuh:152-228) ---------------------------------------------------
void NvbForest::extract(cfa::Mesh2D& mesh, std::vector<double>& u, std::vector<double>& v, std::vector<double>& p,
                        std::vector<double>& uo, std::vector<double>& vo) {
    std::vector<std::pair<uint32_t, int>> keyed;
    for (int e = 0; e < (int)el_.size(); ++e) {
        if (!leaf(e)) continue;
        const int* w = el_[e].v;
        const double xc = (x_[w[0]] + x_[w[1]] + x_[w[2]]) / 3.0, yc = (y_[w[0]] + y_[w[1]] + y_[w[2]]) / 3.0;
        keyed.push_back({morton(xc / Sim::L_CHANNEL, yc / Sim::H_CHANNEL), e});
    }
    std::sort(keyed.begin(), keyed.end());   // by key, ties by slot
    leaves_.resize(keyed.size());
    for (size_t i = 0; i < keyed.size(); ++i) leaves_[i] = keyed[i].second;
    std::vector<int> compact(x_.size(), -1);
    nodes_.clear();
    for (int e : leaves_)
        for (int k = 0; k < 3; ++k)
            if (compact[el_[e].v[k]] < 0) {
                compact[el_[e].v[k]] = (int)nodes_.size();
                nodes_.push_back(el_[e].v[k]);
            }
    const int nn = (int)nodes_.size();
    mesh = cfa::Mesh2D();
    mesh.num_nodes = nn;
    mesh.num_elems = (int)leaves_.size();
    mesh.x.resize(nn);  mesh.y.resize(nn);  mesh.bc_tag.resize(nn);
    u.resize(nn);  v.resize(nn);  p.resize(nn);  uo.resize(nn);  vo.resize(nn);
    for (int i = 0; i < nn; ++i) {
        const int s = nodes_[i];
        mesh.x[i] = x_[s];  mesh.y[i] = y_[s];  mesh.bc_tag[i] = tag_[s];
        u[i] = u_[s];  v[i] = v_[s];  p[i] = p_[s];  uo[i] = uo_[s];  vo[i] = vo_[s];
    }
    mesh.elems.resize(leaves_.size());
    for (size_t i = 0; i < leaves_.size(); ++i) {
        const El& T = el_[leaves_[i]];
        for (int k = 0; k < 3; ++k) mesh.elems[i].v[k] = compact[T.v[k]];
        mesh.elems[i].gen = T.gen;
        mesh.elems[i].level = T.gen / 2;
    }
}

// ---- solution transfer (src/cfa/amr.cuh:96-145, 896-914) ------------------------------------
void NvbForest::scatter_solution(const std::vector<double>& u, const std::vector<double>& v,
                                 const std::vector<double>& p, const std::vector<double>& uo,
                                 const std::vector<double>& vo) {
    for (size_t i = 0; i < nodes_.size(); ++i) {
        const int s = nodes_[i];
        u_[s] = u[i];  v_[s] = v[i];  p_[s] = p[i];  uo_[s] = uo[i];  vo_[s] = vo[i];
    }
}

void NvbForest::retransfer(const std::vector<int>& old_nodes, const std::vector<double>& u,
                           const std::vector<double>& v, const std::vector<double>& p,
                           const std::vector<double>& uo, const std::vector<double>& vo, std::vector<double>& nu,
                           std::vector<double>& nv, std::vector<double>& np, std::vector<double>& nuo,
                           std::vector<double>& nvo) {
    for (size_t j = 0; j < old_nodes.size(); ++j) {
        const int s = old_nodes[j];
        u_[s] = u[j];  v_[s] = v[j];  p_[s] = p[j];  uo_[s] = uo[j];  vo_[s] = vo[j];
    }
    for (const auto& c : created_) midpoint_values(c[0], c[1], c[2]);
    const size_t nn = nodes_.size();
    nu.resize(nn);  nv.resize(nn);  np.resize(nn);  nuo.resize(nn);  nvo.resize(nn);
    for (size_t i = 0; i < nn; ++i) {
        const int s = nodes_[i];
        nu[i] = u_[s];  nv[i] = v_[s];  np[i] = p_[s];  nuo[i] = uo_[s];  nvo[i] = vo_[s];
    }
}

void NvbForest::midpoint_values(int m, int a, int b) {
    u_[m] = 0.5 * (u_[a] + u_[b]);  v_[m] = 0.5 * (v_[a] + v_[b]);  p_[m] = 0.5 * (p_[a] + p_[b]);
    uo_[m] = 0.5 * (uo_[a] + uo_[b]);  vo_[m] = 0.5 * (vo_[a] + vo_[b]);
    if (tag_[m] == cfa::BC_CYLINDER || tag_[m] == cfa::BC_WALL) {
        u_[m] = v_[m] = uo_[m] = vo_[m] = 0.0;
    } else if (tag_[m] == cfa::BC_INLET) {
        u_[m] = uo_[m] = 4.0 * inlet_umax * y_[m] * (Sim::H_CHANNEL - y_[m]) / (Sim::H_CHANNEL * Sim::H_CHANNEL);
        v_[m] = vo_[m] = 0.0;
    }
}

// ---- slots (LIFO recycling, src/cfa/amr.cuh:761-795) ------------------------------------------
int NvbForest::new_vertex() {
    if (!vfree_.empty()) {
        const int s = vfree_.back();
        vfree_.pop_back();
        return s;
    }
    for (auto* a : {&x_, &y_, &u_, &v_, &p_, &uo_, &vo_}) a->push_back(0.0);
    tag_.push_back(cfa::BC_INTERIOR);
    valive_.push_back(0);
    vpar_.push_back({-1, -1});
    vedge_.push_back(0);
    return (int)x_.size() - 1;
}
int NvbForest::new_element() {
    if (!efree_.empty()) {
        const int s = efree_.back();
        efree_.pop_back();
        return s;
    }
    el_.emplace_back();
    return (int)el_.size() - 1;
}
void NvbForest::drop_element(int e) {
    el_[e] = El();
    efree_.push_back(e);
}

// src/cfa/amr.cuh:802-846: midpoint of edge (a, b), created on first use
int NvbForest::midpoint(int a, int b) {
    const uint64_t key = cfa::edge_key(a, b);
    const auto it = mid_.find(key);
    if (it != mid_.end()) return it->second;
    double mx = 0.5 * (x_[a] + x_[b]), my = 0.5 * (y_[a] + y_[b]);
    int tag;
    if (tag_[a] == cfa::BC_CYLINDER && tag_[b] == cfa::BC_CYLINDER) {
        const double th = std::atan2(my - Sim::Y_CYL, mx - Sim::X_CYL);
        mx = Sim::X_CYL + Sim::R_CYL * std::cos(th);
        my = Sim::Y_CYL + Sim::R_CYL * std::sin(th);
        tag = cfa::BC_CYLINDER;
    } else {
        tag = cfa::classify_boundary_node(mx, my);
        if (tag == cfa::BC_CYLINDER) tag = cfa::BC_INTERIOR;
    }
    const int m = new_vertex();
    x_[m] = mx;  y_[m] = my;  tag_[m] = tag;
    midpoint_values(m, a, b);
    created_.push_back({m, a, b});
    valive_[m] = 1;
    vpar_[m] = {-1, -1};
    vedge_[m] = key;
    mid_[key] = m;
    return m;
}

// src/cfa/amr.cuh:852-880: (c, a, b) -> (m, c, a), (m, b, c)
void NvbForest::split(int e) {
    const int c = el_[e].v[0], a = el_[e].v[1], b = el_[e].v[2], gen = el_[e].gen;
    const int m = midpoint(a, b);
    const int c1 = new_element(), c2 = new_element();
    el_[c1].v[0] = m;  el_[c1].v[1] = c;  el_[c1].v[2] = a;
    el_[c2].v[0] = m;  el_[c2].v[1] = b;  el_[c2].v[2] = c;
    for (int k : {c1, c2}) {
        el_[k].gen = gen + 1;
        el_[k].parent = e;
        el_[k].ch[0] = el_[k].ch[1] = -1;
        el_[k].alive = true;
    }
    el_[e].ch[0] = c1;
    el_[e].ch[1] = c2;
    auto& par = vpar_[m];
    if (par[0] < 0) par[0] = e;
    else if (par[1] < 0) par[1] = e;
    else {
        std::fprintf(stderr, "[oracle AMR] edge bisected by more than two elements\n");
        std::exit(EXIT_FAILURE);
    }
}

// src/cfa/amr.cuh:476-508 and 941-969: rounds of bisecting every leaf still below its target
// generation (ascending slot; children inherit the target), each followed by closure passes that
// bisect every hanging leaf (ascending slot) until none is left
void NvbForest::refine(std::vector<int>& target, AmrCounts& st) {
    auto hanging = [&](int e) {
        const int* w = el_[e].v;
        for (int k = 0; k < 3; ++k)
            if (mid_.count(cfa::edge_key(w[k], w[(k + 1) % 3]))) return true;
        return false;
    };
    pass_n_.clear();
    while (true) {
        std::vector<int> work;
        for (int e = 0; e < (int)el_.size(); ++e)
            if (leaf(e) && e < (int)target.size() && target[e] > el_[e].gen) work.push_back(e);
        if (work.empty()) break;
        pass_n_.push_back((int)work.size());
        for (int e : work) {
            split(e);
            ++st.bisections;
            target.resize(el_.size(), -1);
            target[el_[e].ch[0]] = target[e];
            target[el_[e].ch[1]] = target[e];
        }
        while (true) {
            std::vector<int> h;
            for (int e = 0; e < (int)el_.size(); ++e)
                if (leaf(e) && hanging(e)) h.push_back(e);
            if (h.empty()) break;
            pass_n_.push_back(-(int)h.size());
            for (int e : h) {
                split(e);
                ++st.closure_bisections;
                target.resize(el_.size(), -1);
                target[el_[e].ch[0]] = target[el_[e].ch[1]] = -1;
            }
        }
    }
}

AmrCounts NvbForest::refine_uniform(int levels) {
    AmrCounts st;
    st.leaves_before = (int)leaves_.size();
    created_.clear();
    std::vector<int> target(el_.size(), -1);
    for (int e : leaves_) target[e] = el_[e].gen + 2 * levels;
    refine(target, st);
    for (int e = 0; e < (int)el_.size(); ++e) st.leaves_after += leaf(e) ? 1 : 0;
    return st;
}

// ---- one adaptation (src/cfa/amr.cuh:238-455) -------------------------------------------------
AmrCounts NvbForest::adapt(const std::vector<double>& eta, const AmrParams& prm) {
    AmrCounts st;
    created_.clear();
    const int L = (int)leaves_.size(), max_gen = 2 * prm.max_level;
    const size_t NE = el_.size();
    st.leaves_before = L;
    const bool use_v = prm.tol_vort > 0.0, use_i = prm.tol_interp > 0.0;

    // 1. indicators normalised by their tolerances, on forest slots
    std::vector<double> av(NE, 0.0), ai(NE, 0.0);
    for (int i = 0; i < L; ++i) {
        const int e = leaves_[i];
        double w = 1.0;
        if (prm.tol_radius > 0.0) {
            const int* t = el_[e].v;
            const double xc = (x_[t[0]] + x_[t[1]] + x_[t[2]]) / 3.0 - Sim::X_CYL;
            const double yc = (y_[t[0]] + y_[t[1]] + y_[t[2]]) / 3.0 - Sim::Y_CYL;
            const double d = std::sqrt(xc * xc + yc * yc);
            if (d > prm.tol_radius) w = std::pow(d / prm.tol_radius, prm.tol_power);
        }
        if (use_v) av[e] = eta[i] / (prm.tol_vort * w);
        if (use_i) ai[e] = eta[L + i] / (prm.tol_interp * w);
    }

    // edge neighbours of the leaves (compact numbering): nbr[3 i + k] across the edge opposite v[k]
    std::vector<int> pos(NE, -1), nbr(3 * (size_t)L, -1);
    for (int i = 0; i < L; ++i) pos[leaves_[i]] = i;
    {
        std::map<uint64_t, std::vector<std::pair<int, int>>> by_edge;
        for (int i = 0; i < L; ++i) {
            const int* t = el_[leaves_[i]].v;
            for (int k = 0; k < 3; ++k) by_edge[cfa::edge_key(t[(k + 1) % 3], t[(k + 2) % 3])].push_back({i, k});
        }
        for (const auto& kv : by_edge)
            if (kv.second.size() == 2) {
                nbr[3 * kv.second[0].first + kv.second[0].second] = kv.second[1].first;
                nbr[3 * kv.second[1].first + kv.second[1].second] = kv.second[0].first;
            }
    }

    // 2. advective look-ahead (src/cfa/amr.cuh:672-759)
    if (prm.lookahead_time > 0.0) {
        const double horizon = prm.lookahead_time, qv = 1.0, qi = 2.0;
        const double cell = 0.01;
        const int bnx = (int)std::ceil(Sim::L_CHANNEL / cell), bny = (int)std::ceil(Sim::H_CHANNEL / cell);
        std::vector<std::vector<int>> bucket;   // built on first use
        auto clampi = [](int v, int lo, int hi) { return std::max(lo, std::min(hi, v)); };
        auto locate_bucket = [&](double px, double py) {
            if (px < 0.0 || px > Sim::L_CHANNEL || py < 0.0 || py > Sim::H_CHANNEL) return -1;
            if (bucket.empty()) {
                bucket.resize((size_t)bnx * bny);
                for (int e : leaves_) {
                    const int* t = el_[e].v;
                    const double xmin = std::min({x_[t[0]], x_[t[1]], x_[t[2]]}), xmax = std::max({x_[t[0]], x_[t[1]], x_[t[2]]});
                    const double ymin = std::min({y_[t[0]], y_[t[1]], y_[t[2]]}), ymax = std::max({y_[t[0]], y_[t[1]], y_[t[2]]});
                    const int i0 = clampi((int)(xmin / cell), 0, bnx - 1), i1 = clampi((int)(xmax / cell), 0, bnx - 1);
                    const int j0 = clampi((int)(ymin / cell), 0, bny - 1), j1 = clampi((int)(ymax / cell), 0, bny - 1);
                    for (int j = j0; j <= j1; ++j)
                        for (int i = i0; i <= i1; ++i) bucket[(size_t)j * bnx + i].push_back(e);
                }
            }
            const int b = clampi((int)(py / cell), 0, bny - 1) * bnx + clampi((int)(px / cell), 0, bnx - 1);
            double l[3];
            for (int e : bucket[b])
                if (inside(e, px, py, l)) return e;
            return -1;
        };
        auto locate = [&](int e, double px, double py) {
            double l[3];
            for (int hop = 0; hop < 256; ++hop) {
                if (inside(e, px, py, l)) return e;
                const int k = (l[0] < l[1]) ? (l[0] < l[2] ? 0 : 2) : (l[1] < l[2] ? 1 : 2);
                const int nb = nbr[3 * pos[e] + k];
                if (nb < 0) break;
                e = leaves_[nb];
            }
            return locate_bucket(px, py);
        };
        std::vector<double> sv(L), si(L);
        for (int i = 0; i < L; ++i) { sv[i] = av[leaves_[i]]; si[i] = ai[leaves_[i]]; }
        for (int i = 0; i < L; ++i) {
            const bool act_v = use_v && sv[i] > 1.0, act_i = use_i && si[i] > 1.0;
            if (!act_v && !act_i) continue;
            const int e = leaves_[i];
            const int* t = el_[e].v;
            const double xc = (x_[t[0]] + x_[t[1]] + x_[t[2]]) / 3.0, yc = (y_[t[0]] + y_[t[1]] + y_[t[2]]) / 3.0;
            const double uc = (u_[t[0]] + u_[t[1]] + u_[t[2]]) / 3.0, vc = (v_[t[0]] + v_[t[1]] + v_[t[2]]) / 3.0;
            const double hK = std::sqrt(2.0 * area(e));
            const double dist = std::hypot(uc, vc) * horizon;
            if (dist < 0.5 * hK) continue;
            const int n = std::min(64, (int)std::ceil(dist / (0.5 * hK)));
            int last = e;
            for (int s = 1; s <= n; ++s) {
                const double f = horizon * (double)s / n;
                const double px = xc + uc * f, py = yc + vc * f;
                if (px < 0.0 || px > Sim::L_CHANNEL || py < 0.0 || py > Sim::H_CHANNEL) break;
                const int tt = locate(last, px, py);
                if (tt < 0) break;
                if (tt == last) continue;
                last = tt;
                const double hT = std::sqrt(2.0 * area(tt));
                if (act_v) av[tt] = std::max(av[tt], sv[i] * std::pow(hT / hK, qv));
                if (act_i) ai[tt] = std::max(ai[tt], si[i] * std::pow(hT / hK, qi));
            }
        }
    }

    auto score = [&](int e) { return std::max(av[e], ai[e]); };
    auto parent_score = [&](int e) { return std::max(std::sqrt(2.0) * av[e], 2.0 * ai[e]); };

    // 3. refinement marks within the element budget, largest score first (ties in leaf order)
    std::vector<int> target(NE, -1);
    std::vector<char> refine_mark(NE, 0);
    long long predicted = L;
    auto mark = [&](int e, int tgt) {
        const int gen = el_[e].gen;
        tgt = std::min(max_gen, tgt);
        if (tgt <= gen) return;
        const long long added = (1LL << (tgt - gen)) - 1;
        if (predicted + added > prm.max_elems) return;
        predicted += added;
        refine_mark[e] = 1;
        target[e] = tgt;
        ++st.marked_refine;
    };
    std::vector<int> order;
    for (int e : leaves_)
        if (score(e) > 1.0) order.push_back(e);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return score(a) > score(b); });
    for (int e : order) {
        int levels = 1;
        if (av[e] > 1.0) levels = std::max(levels, (int)std::ceil(std::log2(av[e]) / 1.0));
        if (ai[e] > 1.0) levels = std::max(levels, (int)std::ceil(std::log2(ai[e]) / 2.0));
        mark(e, el_[e].gen + 2 * levels);
    }
    if (prm.min_level > 0)
        for (int e : leaves_)
            if (!refine_mark[e] && el_[e].gen < 2 * prm.min_level) mark(e, 2 * prm.min_level);

    // 4. hole filling: two passes, each decided before it marks
    auto planned = [&](int e) { return std::max(el_[e].gen, target[e]) / 2; };
    auto finer = [&](int e) {
        const int lvl = planned(e);
        int c = 0;
        for (int k = 0; k < 3; ++k) {
            const int nb = nbr[3 * pos[e] + k];
            if (nb >= 0 && planned(leaves_[nb]) > lvl) ++c;
        }
        return c;
    };
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fill;
        for (int e : leaves_)
            if (!refine_mark[e] && el_[e].gen < max_gen && score(e) > prm.coarsen_safety && finer(e) >= 2)
                fill.push_back(e);
        for (int e : fill) mark(e, el_[e].gen + 2);
    }

    // 5. buffer layers, then one protected layer
    std::vector<char> vflag(x_.size(), 0);
    auto flag_marked = [&]() {
        std::fill(vflag.begin(), vflag.end(), 0);
        for (int e : leaves_)
            if (refine_mark[e])
                for (int k = 0; k < 3; ++k) vflag[el_[e].v[k]] = 1;
    };
    auto touches = [&](int e) { const int* t = el_[e].v; return vflag[t[0]] || vflag[t[1]] || vflag[t[2]]; };
    for (int layer = 0; layer < prm.buffer_layers; ++layer) {
        flag_marked();
        for (int e : leaves_)
            if (!refine_mark[e] && el_[e].gen < max_gen && touches(e)) mark(e, el_[e].gen + 2);
    }
    flag_marked();

    // 6. coarsening by collapsing newest vertices (at most two sweeps)
    std::vector<char> coarsen(NE, 0);
    std::vector<int> cand_leaves;
    for (int e : leaves_) {
        if (refine_mark[e] || el_[e].gen <= 2 * prm.min_level || touches(e)) continue;
        if (prm.keep_radius > 0.0) {
            const int* t = el_[e].v;
            const double xc = (x_[t[0]] + x_[t[1]] + x_[t[2]]) / 3.0 - Sim::X_CYL;
            const double yc = (y_[t[0]] + y_[t[1]] + y_[t[2]]) / 3.0 - Sim::Y_CYL;
            if (xc * xc + yc * yc < prm.keep_radius * prm.keep_radius) continue;
        }
        if (parent_score(e) < prm.coarsen_safety && finer(e) < 2) {
            coarsen[e] = 1;
            cand_leaves.push_back(e);
            ++st.marked_coarsen;
        }
    }
    for (int sweep = 0; sweep < 2; ++sweep) {
        std::vector<int> verts;
        for (int e : cand_leaves)
            if (leaf(e) && el_[e].gen > 0) verts.push_back(el_[e].v[0]);
        std::sort(verts.begin(), verts.end());
        verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
        std::vector<int> go;
        for (int m : verts) {
            bool ok = false;
            for (int s = 0; s < 2; ++s) {
                const int pe = vpar_[m][s];
                if (pe < 0) continue;
                const int c0 = el_[pe].ch[0], c1 = el_[pe].ch[1];
                ok = el_[pe].alive && c0 >= 0 && leaf(c0) && leaf(c1) && coarsen[c0] && coarsen[c1] &&
                     std::max(parent_score(c0), parent_score(c1)) < prm.coarsen_safety;
                if (!ok) break;
            }
            if (ok) go.push_back(m);
        }
        if (go.empty()) break;
        cand_leaves.clear();
        for (int m : go) {
            for (int s = 0; s < 2; ++s) {
                const int pe = vpar_[m][s];
                if (pe < 0) continue;
                cand_leaves.push_back(pe);
                const int c0 = el_[pe].ch[0], c1 = el_[pe].ch[1];
                const double mv = std::sqrt(2.0) * std::max(av[c0], av[c1]);
                const double mi = 2.0 * std::max(ai[c0], ai[c1]);
                drop_element(c0);
                drop_element(c1);
                el_[pe].ch[0] = el_[pe].ch[1] = -1;
                coarsen[pe] = 1;
                av[pe] = mv;
                ai[pe] = mi;
            }
            mid_.erase(vedge_[m]);
            valive_[m] = 0;
            vpar_[m] = {-1, -1};
            vfree_.push_back(m);
            ++st.removed_vertices;
        }
    }

    // 7. refinement to the targets with conformity closure
    refine(target, st);
    for (int e = 0; e < (int)el_.size(); ++e) st.leaves_after += leaf(e) ? 1 : 0;
    return st;
}

}  // namespace oracle
