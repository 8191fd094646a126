// NVB refinement forest (host). Restates the decision rules of src/cfa/amr.cuh:238-996; the
// order in which elements are visited, bisected and recycled is part of the result (it fixes
// forest ids, recycled vertex slots and, through extract(), the compact numbering), so it is
// kept exactly: ascending element id for bisection rounds and closure passes, LIFO reuse of
// freed slots, marking by descending score with ties in leaf order.
#include "amr.hpp"

#include <algorithm>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cfa {

AdaptiveMesh::AdaptiveMesh(const Mesh2D& base, const std::vector<double>& u0,
                           const std::vector<double>& v0, const std::vector<double>& p0) {
    x_ = base.x;
    y_ = base.y;
    bc_tag_ = base.bc_tag;
    u_ = u0;  v_ = v0;  p_ = p0;
    uo_ = u0; vo_ = v0;
    const int nv = base.num_nodes;
    v_alive_.assign(nv, 1);
    v_parents_.assign(nv, {-1, -1});
    v_edge_.assign(nv, 0);
    elems_.resize(base.num_elems);
    for (int e = 0; e < base.num_elems; ++e) {
        for (int k = 0; k < 3; ++k) elems_[e].v[k] = base.elems[e].v[k];
        elems_[e].gen = 0;
        elems_[e].alive = true;
    }
}

// ---------------------------------------------------------------------------------------------
// Solution transfer
// ---------------------------------------------------------------------------------------------
void AdaptiveMesh::scatter_solution(const std::vector<double>& h_u, const std::vector<double>& h_v,
                                    const std::vector<double>& h_p, const std::vector<double>& h_uo,
                                    const std::vector<double>& h_vo) {
    for (size_t i = 0; i < vert_ids_.size(); ++i) {
        const int s = vert_ids_[i];
        u_[s] = h_u[i];  v_[s] = h_v[i];  p_[s] = h_p[i];
        uo_[s] = h_uo[i];  vo_[s] = h_vo[i];
    }
}

// Replays the transfer of the last adapt() for a later state on the mesh it started from
// (src/cfa/amr.cuh:127-145): kept vertices take the new values, created midpoints are
// re-interpolated in creation order.
void AdaptiveMesh::retransfer(const std::vector<int>& old_vids, const std::vector<double>& h_u,
                              const std::vector<double>& h_v, const std::vector<double>& h_p,
                              const std::vector<double>& h_uo, const std::vector<double>& h_vo,
                              std::vector<double>& n_u, std::vector<double>& n_v,
                              std::vector<double>& n_p, std::vector<double>& n_uo,
                              std::vector<double>& n_vo) {
    const int no = (int)old_vids.size();
#pragma omp parallel for schedule(static)
    for (int j = 0; j < no; ++j) {
        const int s = old_vids[j];
        u_[s] = h_u[j];  v_[s] = h_v[j];  p_[s] = h_p[j];  uo_[s] = h_uo[j];  vo_[s] = h_vo[j];
    }
    for (const auto& c : created_log_) set_midpoint_values(c[0], c[1], c[2]);
    const int nn = (int)vert_ids_.size();
    n_u.resize(nn);  n_v.resize(nn);  n_p.resize(nn);  n_uo.resize(nn);  n_vo.resize(nn);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nn; ++i) {
        const int s = vert_ids_[i];
        n_u[i] = u_[s];  n_v[i] = v_[s];  n_p[i] = p_[s];  n_uo[i] = uo_[s];  n_vo[i] = vo_[s];
    }
}

// ---------------------------------------------------------------------------------------------
// Extraction: leaves along the Morton curve of their centroids (ties by forest id), vertices by
// first use (src/cfa/amr.cuh:152-228, 307-321).
// ---------------------------------------------------------------------------------------------
uint32_t AdaptiveMesh::spread_bits(uint32_t n) {
    n &= 0x0000ffffu;
    n = (n | (n << 8)) & 0x00ff00ffu;
    n = (n | (n << 4)) & 0x0f0f0f0fu;
    n = (n | (n << 2)) & 0x33333333u;
    n = (n | (n << 1)) & 0x55555555u;
    return n;
}

uint32_t AdaptiveMesh::morton_code(double sx, double sy) {
    auto q16 = [](double s) { return (uint32_t)std::min(65535.0, std::max(0.0, s * 65535.0)); };
    return spread_bits(q16(sx)) | (spread_bits(q16(sy)) << 1);
}

void AdaptiveMesh::extract(Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                           std::vector<double>& h_p, std::vector<double>& h_uo,
                           std::vector<double>& h_vo) {
    dbg_start();
    leaf_ids_.clear();
    for (int e = 0; e < (int)elems_.size(); ++e)
        if (elems_[e].is_leaf()) leaf_ids_.push_back(e);
    dbg_lap(D_COLLECT);

    const long long nl = (long long)leaf_ids_.size();
    std::vector<std::pair<uint32_t, int>> keyed(nl);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nl; ++i) {
        const AmrElem& t = elems_[leaf_ids_[i]];
        const double xc = (x_[t.v[0]] + x_[t.v[1]] + x_[t.v[2]]) / 3.0;
        const double yc = (y_[t.v[0]] + y_[t.v[1]] + y_[t.v[2]]) / 3.0;
        keyed[i] = {morton_code(xc / Sim::L_CHANNEL, yc / Sim::H_CHANNEL), leaf_ids_[i]};
    }
    dbg_lap(D_KEYS);
    // (key, id) is a total order, so any correct sort gives the same permutation
    if (!leaf_sorter || !leaf_sorter(keyed)) std::sort(keyed.begin(), keyed.end());
    for (long long i = 0; i < nl; ++i) leaf_ids_[i] = keyed[i].second;
    dbg_lap(D_SORT);

    std::vector<int> compact(x_.size(), -1);
    vert_ids_.clear();
    for (int e : leaf_ids_)
        for (int k = 0; k < 3; ++k) {
            const int vid = elems_[e].v[k];
            if (compact[vid] < 0) {
                compact[vid] = (int)vert_ids_.size();
                vert_ids_.push_back(vid);
            }
        }
    dbg_lap(D_COMPACT);

    const int nn = (int)vert_ids_.size();
    mesh.x.resize(nn);  mesh.y.resize(nn);  mesh.bc_tag.resize(nn);
    h_u.resize(nn);  h_v.resize(nn);  h_p.resize(nn);  h_uo.resize(nn);  h_vo.resize(nn);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nn; ++i) {
        const int s = vert_ids_[i];
        mesh.x[i] = x_[s];  mesh.y[i] = y_[s];  mesh.bc_tag[i] = bc_tag_[s];
        h_u[i] = u_[s];  h_v[i] = v_[s];  h_p[i] = p_[s];  h_uo[i] = uo_[s];  h_vo[i] = vo_[s];
    }
    mesh.elems.resize(nl);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < nl; ++i) {
        const AmrElem& t = elems_[leaf_ids_[i]];
        Triangle tri;
        for (int k = 0; k < 3; ++k) tri.v[k] = compact[t.v[k]];
        tri.gen = t.gen;
        tri.level = t.gen / 2;
        mesh.elems[i] = tri;
    }
    mesh.num_nodes = nn;
    mesh.num_elems = (int)nl;
    dbg_lap(D_FILL);
}

// ---------------------------------------------------------------------------------------------
// Adaptation cycle (src/cfa/amr.cuh:238-455)
// ---------------------------------------------------------------------------------------------
AmrStats AdaptiveMesh::adapt(const std::vector<double>& eta, const AmrParams& prm) {
    const auto t_start = Clock::now();
    AmrStats st;
    created_log_.clear();
    const int max_gen = 2 * prm.max_level;
    const int L = (int)leaf_ids_.size();
    st.leaves_before = L;
    constexpr double Q_VORT = 1.0, Q_INTERP = 2.0;
    const bool use_v = prm.tol_vort > 0.0, use_i = prm.tol_interp > 0.0;

    // 1. indicators normalised by their tolerances (optionally growing with the distance)
    std::vector<double> av(elems_.size(), 0.0), ai(elems_.size(), 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L; ++i) {
        double w = 1.0;
        if (prm.tol_radius > 0.0) {
            const AmrElem& t = elems_[leaf_ids_[i]];
            const double xc = (x_[t.v[0]] + x_[t.v[1]] + x_[t.v[2]]) / 3.0 - Sim::X_CYL;
            const double yc = (y_[t.v[0]] + y_[t.v[1]] + y_[t.v[2]]) / 3.0 - Sim::Y_CYL;
            const double d = std::sqrt(xc * xc + yc * yc);
            if (d > prm.tol_radius) w = std::pow(d / prm.tol_radius, prm.tol_power);
        }
        if (use_v) av[leaf_ids_[i]] = eta[i] / (prm.tol_vort * w);
        if (use_i) ai[leaf_ids_[i]] = eta[L + i] / (prm.tol_interp * w);
    }

    const auto t_look = Clock::now();
    dbg_start();
    build_leaf_adjacency();
    dbg_lap(D_ADJ);
    buckets_built_.store(false);

    // 2. advective look-ahead
    if (prm.lookahead_time > 0.0)
        apply_lookahead(use_v ? &av : nullptr, use_i ? &ai : nullptr, prm.lookahead_time, Q_VORT, Q_INTERP);
    st.ms_lookahead = ms_since(t_look);

    auto score = [&](int e) { return std::max(av[e], ai[e]); };
    auto parent_score = [&](int e) { return std::max(std::sqrt(2.0) * av[e], 2.0 * ai[e]); };

    // 3. tolerance marking, largest score first (ties in leaf order), within the element budget
    std::vector<std::pair<double, int>> ranked;
    for (int e : leaf_ids_)
        if (score(e) > 1.0) ranked.push_back({score(e), e});
    std::stable_sort(ranked.begin(), ranked.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first > b.first; });

    std::vector<int> target(elems_.size(), -1);
    std::vector<char> refine(elems_.size(), 0);
    long long predicted = L;
    auto mark = [&](int e, int tgt) {
        const int gen = elems_[e].gen;
        tgt = std::min(max_gen, tgt);
        if (tgt <= gen) return false;
        const long long added = (1LL << (tgt - gen)) - 1;
        if (predicted + added > prm.max_elems) return false;
        predicted += added;
        refine[e] = 1;
        target[e] = tgt;
        ++st.marked_refine;
        return true;
    };
    for (const auto& se : ranked) {
        const int e = se.second;
        if (score(e) <= 1.0) break;
        int levels = 1;
        if (av[e] > 1.0) levels = std::max(levels, (int)std::ceil(std::log2(av[e]) / Q_VORT));
        if (ai[e] > 1.0) levels = std::max(levels, (int)std::ceil(std::log2(ai[e]) / Q_INTERP));
        mark(e, elems_[e].gen + 2 * levels);
    }
    if (prm.min_level > 0)
        for (int e : leaf_ids_)
            if (!refine[e] && elems_[e].gen < 2 * prm.min_level) mark(e, 2 * prm.min_level);

    st.ms_mark = ms_since(t_start) - st.ms_lookahead;
    auto t_phase = Clock::now();

    // 4. hole filling (two passes): leaves coarser than two edge neighbours and not far below
    //    tolerance get one more level
    auto planned_level = [&](int e) { return std::max(elems_[e].gen, target[e]) / 2; };
    auto finer_neighbours = [&](int e) {
        const int* nb = &leaf_nbr_[3 * leaf_pos_[e]];
        const int lvl = planned_level(e);
        int count = 0;
        for (int k = 0; k < 3; ++k)
            if (nb[k] >= 0 && planned_level(leaf_ids_[nb[k]]) > lvl) ++count;
        return count;
    };
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int> fill;
        for (int e : leaf_ids_) {
            if (refine[e] || elems_[e].gen >= max_gen) continue;
            if (score(e) > prm.coarsen_safety && finer_neighbours(e) >= 2) fill.push_back(e);
        }
        for (int e : fill) mark(e, elems_[e].gen + 2);
    }

    // 5. buffer layers around marked leaves (vertex neighbours)
    std::vector<char> vflag(x_.size(), 0);
    for (int layer = 0; layer < prm.buffer_layers; ++layer) {
        std::fill(vflag.begin(), vflag.end(), 0);
        for (int e : leaf_ids_)
            if (refine[e])
                for (int k = 0; k < 3; ++k) vflag[elems_[e].v[k]] = 1;
        for (int e : leaf_ids_) {
            const AmrElem& t = elems_[e];
            if (refine[e] || t.gen >= max_gen) continue;
            if (vflag[t.v[0]] || vflag[t.v[1]] || vflag[t.v[2]]) mark(e, t.gen + 2);
        }
    }
    // one further layer is protected from coarsening
    std::fill(vflag.begin(), vflag.end(), 0);
    for (int e : leaf_ids_)
        if (refine[e])
            for (int k = 0; k < 3; ++k) vflag[elems_[e].v[k]] = 1;

    st.ms_neighbours = ms_since(t_phase);
    t_phase = Clock::now();

    // 6. coarsening: newest vertices whose merged parents are predicted below coarsen_safety
    std::vector<char> coarsen(elems_.size(), 0);
    std::vector<int> coarsen_leaves;
    for (int e : leaf_ids_) {
        const AmrElem& t = elems_[e];
        if (refine[e] || t.gen <= 2 * prm.min_level) continue;
        if (vflag[t.v[0]] || vflag[t.v[1]] || vflag[t.v[2]]) continue;
        if (prm.keep_radius > 0.0) {
            const double xc = (x_[t.v[0]] + x_[t.v[1]] + x_[t.v[2]]) / 3.0 - Sim::X_CYL;
            const double yc = (y_[t.v[0]] + y_[t.v[1]] + y_[t.v[2]]) / 3.0 - Sim::Y_CYL;
            if (xc * xc + yc * yc < prm.keep_radius * prm.keep_radius) continue;
        }
        if (parent_score(e) < prm.coarsen_safety && finer_neighbours(e) < 2) {
            coarsen[e] = 1;
            coarsen_leaves.push_back(e);
            ++st.marked_coarsen;
        }
    }
    std::vector<char> seen(x_.size(), 0);
    for (int sweep = 0; sweep < 2; ++sweep) {
        std::vector<int> candidates, examined;
        for (int e : coarsen_leaves) {
            if (!elems_[e].is_leaf() || elems_[e].gen == 0) continue;
            const int m = elems_[e].v[0];
            if (seen[m]) continue;
            seen[m] = 1;
            examined.push_back(m);
            bool ok = false;
            for (int s = 0; s < 2; ++s) {
                const int pe = v_parents_[m][s];
                if (pe < 0) continue;
                const AmrElem& par = elems_[pe];
                const int c0 = par.child[0], c1 = par.child[1];
                const bool removable = par.alive && c0 >= 0 && elems_[c0].is_leaf() &&
                                       elems_[c1].is_leaf() && coarsen[c0] && coarsen[c1] &&
                                       std::max(parent_score(c0), parent_score(c1)) < prm.coarsen_safety;
                if (!removable) { ok = false; break; }
                ok = true;
            }
            if (ok) candidates.push_back(m);
        }
        for (int m : examined) seen[m] = 0;
        std::sort(candidates.begin(), candidates.end());
        coarsen_leaves.clear();
        for (int m : candidates) {
            for (int pe : v_parents_[m])
                if (pe >= 0) coarsen_leaves.push_back(pe);
            collapse_vertex(m, coarsen, av, ai);
            ++st.removed_vertices;
        }
        if (candidates.empty()) break;
    }
    st.ms_coarsen = ms_since(t_phase);
    t_phase = Clock::now();

    // 7. refinement to the target generations with conformity closure
    refine_to_targets(target, st);
    st.ms_refine = ms_since(t_phase);

    int leaves = 0;
    for (const AmrElem& t : elems_) leaves += t.is_leaf() ? 1 : 0;
    st.leaves_after = leaves;
    st.elapsed_ms = ms_since(t_start);
    return st;
}

AmrStats AdaptiveMesh::refine_uniform(int levels) {
    AmrStats st;
    st.leaves_before = (int)leaf_ids_.size();
    std::vector<int> target(elems_.size(), -1);
    for (int e : leaf_ids_) target[e] = elems_[e].gen + 2 * levels;
    refine_to_targets(target, st);
    int leaves = 0;
    for (const AmrElem& t : elems_) leaves += t.is_leaf() ? 1 : 0;
    st.leaves_after = leaves;
    return st;
}

// Rounds: bisect every working leaf (ascending id), close the mesh, continue with the children
// that still carry a target (src/cfa/amr.cuh:478-508).
void AdaptiveMesh::refine_to_targets(std::vector<int>& target, AmrStats& st) {
    build_vertex_leaves();
    std::vector<int> work, children, next;
    std::vector<std::pair<int, int>> edges;
    for (int e : leaf_ids_)
        if (target[e] > elems_[e].gen) work.push_back(e);
    std::sort(work.begin(), work.end());
    while (!work.empty()) {
        children.clear();
        edges.clear();
        for (int e : work) {
            bisect(e, &children, &edges);
            ++st.bisections;
            target.resize(elems_.size(), -1);
            target[elems_[e].child[0]] = target[e];
            target[elems_[e].child[1]] = target[e];
        }
        st.closure_bisections += make_conforming(children, edges);
        target.resize(elems_.size(), -1);
        next.clear();
        for (int e : work)
            for (int c : elems_[e].child)
                if (elems_[c].is_leaf() && target[c] > elems_[c].gen) next.push_back(c);
        std::sort(next.begin(), next.end());
        work.swap(next);
    }
}

// ---------------------------------------------------------------------------------------------
// Leaf adjacency, point location and look-ahead (src/cfa/amr.cuh:561-759)
// ---------------------------------------------------------------------------------------------
double AdaptiveMesh::leaf_area(int e) const {
    const AmrElem& t = elems_[e];
    return 0.5 * std::fabs((x_[t.v[1]] - x_[t.v[0]]) * (y_[t.v[2]] - y_[t.v[0]]) -
                           (x_[t.v[2]] - x_[t.v[0]]) * (y_[t.v[1]] - y_[t.v[0]]));
}

void AdaptiveMesh::build_leaf_adjacency() {
    const int L = (int)leaf_ids_.size();
    leaf_pos_.assign(elems_.size(), -1);
    for (int i = 0; i < L; ++i) leaf_pos_[leaf_ids_[i]] = i;
    inc_ptr_.assign(x_.size() + 1, 0);
    for (int i = 0; i < L; ++i)
        for (int k = 0; k < 3; ++k) inc_ptr_[elems_[leaf_ids_[i]].v[k] + 1]++;
    for (size_t i = 0; i < x_.size(); ++i) inc_ptr_[i + 1] += inc_ptr_[i];
    inc_list_.resize(inc_ptr_[x_.size()]);
    {
        std::vector<int> pos(inc_ptr_.begin(), inc_ptr_.end() - 1);
        for (int i = 0; i < L; ++i)
            for (int k = 0; k < 3; ++k) inc_list_[pos[elems_[leaf_ids_[i]].v[k]]++] = i;
    }
    // neighbour across the edge opposite v[k]: the other leaf containing both ends (unique on a
    // conforming mesh, so the search order does not matter)
    leaf_nbr_.assign((size_t)3 * L, -1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L; ++i) {
        const AmrElem& t = elems_[leaf_ids_[i]];
        for (int k = 0; k < 3; ++k) {
            const int a = t.v[(k + 1) % 3], b = t.v[(k + 2) % 3];
            for (int s = inc_ptr_[a]; s < inc_ptr_[a + 1]; ++s) {
                const int j = inc_list_[s];
                if (j == i) continue;
                const int* w = elems_[leaf_ids_[j]].v;
                if (w[0] == b || w[1] == b || w[2] == b) { leaf_nbr_[(size_t)3 * i + k] = j; break; }
            }
        }
    }
}

bool AdaptiveMesh::barycentric(int e, double px, double py, double l[3]) const {
    const AmrElem& t = elems_[e];
    const double x0 = x_[t.v[0]], y0 = y_[t.v[0]];
    const double d = (x_[t.v[1]] - x0) * (y_[t.v[2]] - y0) - (x_[t.v[2]] - x0) * (y_[t.v[1]] - y0);
    l[1] = ((px - x0) * (y_[t.v[2]] - y0) - (x_[t.v[2]] - x0) * (py - y0)) / d;
    l[2] = ((x_[t.v[1]] - x0) * (py - y0) - (px - x0) * (y_[t.v[1]] - y0)) / d;
    l[0] = 1.0 - l[1] - l[2];
    return l[1] >= -1.0e-12 && l[2] >= -1.0e-12 && l[1] + l[2] <= 1.0 + 1.0e-12;
}

void AdaptiveMesh::build_buckets() {
    const double cell = BUCKET_CELL;
    bucket_nx_ = (int)std::ceil(Sim::L_CHANNEL / cell);
    bucket_ny_ = (int)std::ceil(Sim::H_CHANNEL / cell);
    const int nx = bucket_nx_, ny = bucket_ny_;
    auto clampi = [](int v, int lo, int hi) { return std::max(lo, std::min(hi, v)); };
    auto range = [&](int e, int& i0, int& i1, int& j0, int& j1) {
        const AmrElem& t = elems_[e];
        const double xmin = std::min({x_[t.v[0]], x_[t.v[1]], x_[t.v[2]]});
        const double xmax = std::max({x_[t.v[0]], x_[t.v[1]], x_[t.v[2]]});
        const double ymin = std::min({y_[t.v[0]], y_[t.v[1]], y_[t.v[2]]});
        const double ymax = std::max({y_[t.v[0]], y_[t.v[1]], y_[t.v[2]]});
        i0 = clampi((int)(xmin / cell), 0, nx - 1);  i1 = clampi((int)(xmax / cell), 0, nx - 1);
        j0 = clampi((int)(ymin / cell), 0, ny - 1);  j1 = clampi((int)(ymax / cell), 0, ny - 1);
    };
    bucket_ptr_.assign((size_t)nx * ny + 1, 0);
    for (int e : leaf_ids_) {
        int i0, i1, j0, j1;
        range(e, i0, i1, j0, j1);
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) bucket_ptr_[(size_t)j * nx + i + 1]++;
    }
    for (int b = 0; b < nx * ny; ++b) bucket_ptr_[b + 1] += bucket_ptr_[b];
    bucket_items_.resize(bucket_ptr_[(size_t)nx * ny]);
    std::vector<int> pos(bucket_ptr_.begin(), bucket_ptr_.end() - 1);
    for (int e : leaf_ids_) {
        int i0, i1, j0, j1;
        range(e, i0, i1, j0, j1);
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) bucket_items_[pos[(size_t)j * nx + i]++] = e;
    }
}

int AdaptiveMesh::locate_bucket(double px, double py) {
    if (px < 0.0 || px > Sim::L_CHANNEL || py < 0.0 || py > Sim::H_CHANNEL) return -1;
    if (!buckets_built_.load(std::memory_order_acquire)) {
        std::lock_guard<std::mutex> lock(buckets_mu_);
        if (!buckets_built_.load(std::memory_order_relaxed)) {
            build_buckets();
            buckets_built_.store(true, std::memory_order_release);
        }
    }
    auto clampi = [](int v, int lo, int hi) { return std::max(lo, std::min(hi, v)); };
    const int b = clampi((int)(py / BUCKET_CELL), 0, bucket_ny_ - 1) * bucket_nx_ +
                  clampi((int)(px / BUCKET_CELL), 0, bucket_nx_ - 1);
    double l[3];
    for (int k = bucket_ptr_[b]; k < bucket_ptr_[b + 1]; ++k)
        if (barycentric(bucket_items_[k], px, py, l)) return bucket_items_[k];
    return -1;
}

int AdaptiveMesh::locate_from(int start, double px, double py) {
    int e = start;
    double l[3];
    for (int hop = 0; hop < 256; ++hop) {
        if (barycentric(e, px, py, l)) return e;
        const int k = (l[0] < l[1]) ? (l[0] < l[2] ? 0 : 2) : (l[1] < l[2] ? 1 : 2);
        const int nb = leaf_nbr_[(size_t)3 * leaf_pos_[e] + k];
        if (nb < 0) break;
        e = leaf_ids_[nb];
    }
    return locate_bucket(px, py);
}

namespace {
// slot = max(slot, val) with a compare-and-swap loop; the final value is the maximum over all
// writers whatever the interleaving.
inline void atomic_max(double* slot, double val) {
    double cur;
    __atomic_load(slot, &cur, __ATOMIC_RELAXED);
    while (cur < val) {
        if (__atomic_compare_exchange(slot, &cur, &val, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return;
    }
}
}  // namespace

void AdaptiveMesh::apply_lookahead(std::vector<double>* av, std::vector<double>* ai,
                                   double horizon, double q_v, double q_i) {
    dbg_start();
    const int L = (int)leaf_ids_.size();
    std::vector<double> src_v(av ? L : 0), src_i(ai ? L : 0);
    for (int i = 0; i < L; ++i) {
        if (av) src_v[i] = (*av)[leaf_ids_[i]];
        if (ai) src_i[i] = (*ai)[leaf_ids_[i]];
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < L; ++i) {
        const bool act_v = av && src_v[i] > 1.0;
        const bool act_i = ai && src_i[i] > 1.0;
        if (!act_v && !act_i) continue;
        const int e = leaf_ids_[i];
        const AmrElem& t = elems_[e];
        const double xc = (x_[t.v[0]] + x_[t.v[1]] + x_[t.v[2]]) / 3.0;
        const double yc = (y_[t.v[0]] + y_[t.v[1]] + y_[t.v[2]]) / 3.0;
        const double uc = (u_[t.v[0]] + u_[t.v[1]] + u_[t.v[2]]) / 3.0;
        const double vc = (v_[t.v[0]] + v_[t.v[1]] + v_[t.v[2]]) / 3.0;
        const double hK = std::sqrt(2.0 * leaf_area(e));
        const double dist = std::hypot(uc, vc) * horizon;
        if (dist < 0.5 * hK) continue;
        const int n = std::min(64, (int)std::ceil(dist / (0.5 * hK)));
        int last = e;
        for (int s = 1; s <= n; ++s) {
            const double f = horizon * (double)s / n;
            const double px = xc + uc * f, py = yc + vc * f;
            if (px < 0.0 || px > Sim::L_CHANNEL || py < 0.0 || py > Sim::H_CHANNEL) break;
            const int tl = locate_from(last, px, py);
            if (tl < 0) break;
            if (tl == last) continue;
            last = tl;
            const double hT = std::sqrt(2.0 * leaf_area(tl));
            if (act_v) atomic_max(&(*av)[tl], src_v[i] * std::pow(hT / hK, q_v));
            if (act_i) atomic_max(&(*ai)[tl], src_i[i] * std::pow(hT / hK, q_i));
        }
    }
    dbg_lap(D_TRACE);
}

// ---------------------------------------------------------------------------------------------
// Forest storage and bisection (src/cfa/amr.cuh:761-878)
// ---------------------------------------------------------------------------------------------
int AdaptiveMesh::alloc_vertex() {
    if (!free_verts_.empty()) {
        const int id = free_verts_.back();
        free_verts_.pop_back();
        return id;
    }
    x_.push_back(0.0);  y_.push_back(0.0);
    u_.push_back(0.0);  v_.push_back(0.0);  p_.push_back(0.0);
    uo_.push_back(0.0); vo_.push_back(0.0);
    bc_tag_.push_back(BC_INTERIOR);
    v_alive_.push_back(0);
    v_parents_.push_back({-1, -1});
    v_edge_.push_back(0);
    return (int)x_.size() - 1;
}

int AdaptiveMesh::alloc_elem() {
    if (!free_elems_.empty()) {
        const int id = free_elems_.back();
        free_elems_.pop_back();
        return id;
    }
    elems_.emplace_back();
    return (int)elems_.size() - 1;
}

void AdaptiveMesh::free_elem(int e) {
    elems_[e] = AmrElem();
    free_elems_.push_back(e);
}

int AdaptiveMesh::get_or_create_midpoint(int a, int b) {
    const uint64_t key = edge_key(a, b);
    auto it = edge_mid_.find(key);
    if (it != edge_mid_.end()) return it->second;
    double mx = 0.5 * (x_[a] + x_[b]);
    double my = 0.5 * (y_[a] + y_[b]);
    int tag;
    if (bc_tag_[a] == BC_CYLINDER && bc_tag_[b] == BC_CYLINDER) {
        // cylinder nodes are joined only by boundary edges: project onto the circle
        const double th = std::atan2(my - Sim::Y_CYL, mx - Sim::X_CYL);
        mx = Sim::X_CYL + Sim::R_CYL * std::cos(th);
        my = Sim::Y_CYL + Sim::R_CYL * std::sin(th);
        tag = BC_CYLINDER;
    } else {
        // straight sides by coordinates; a cylinder tag here would be a boundary-layer false hit
        tag = classify_boundary_node(mx, my);
        if (tag == BC_CYLINDER) tag = BC_INTERIOR;
    }
    const int m = alloc_vertex();
    x_[m] = mx;  y_[m] = my;  bc_tag_[m] = tag;
    set_midpoint_values(m, a, b);
    created_log_.push_back({m, a, b});
    v_alive_[m] = 1;
    v_parents_[m] = {-1, -1};
    v_edge_[m] = key;
    edge_mid_[key] = m;
    return m;
}

void AdaptiveMesh::set_midpoint_values(int m, int a, int b) {
    u_[m] = 0.5 * (u_[a] + u_[b]);     v_[m] = 0.5 * (v_[a] + v_[b]);     p_[m] = 0.5 * (p_[a] + p_[b]);
    uo_[m] = 0.5 * (uo_[a] + uo_[b]);  vo_[m] = 0.5 * (vo_[a] + vo_[b]);
    if (bc_tag_[m] == BC_CYLINDER || bc_tag_[m] == BC_WALL) {
        u_[m] = v_[m] = uo_[m] = vo_[m] = 0.0;
    } else if (bc_tag_[m] == BC_INLET) {
        u_[m] = uo_[m] = 4.0 * inlet_umax * y_[m] * (Sim::H_CHANNEL - y_[m]) / (Sim::H_CHANNEL * Sim::H_CHANNEL);
        v_[m] = vo_[m] = 0.0;
    }
}

// (c, a, b) -> (m, c, a), (m, b, c): m newest vertex of both children.
void AdaptiveMesh::bisect(int e, std::vector<int>* new_children,
                          std::vector<std::pair<int, int>>* new_edges) {
    const int c = elems_[e].v[0], a = elems_[e].v[1], b = elems_[e].v[2];
    const int gen = elems_[e].gen;
    if (new_edges && !edge_mid_.count(edge_key(a, b))) new_edges->push_back({a, b});
    const int m = get_or_create_midpoint(a, b);
    const int k1 = alloc_elem();
    const int k2 = alloc_elem();
    AmrElem& t1 = elems_[k1];
    t1.v[0] = m;  t1.v[1] = c;  t1.v[2] = a;
    t1.gen = gen + 1;  t1.parent = e;  t1.alive = true;
    AmrElem& t2 = elems_[k2];
    t2.v[0] = m;  t2.v[1] = b;  t2.v[2] = c;
    t2.gen = gen + 1;  t2.parent = e;  t2.alive = true;
    elems_[e].child[0] = k1;
    elems_[e].child[1] = k2;
    if (new_children) { new_children->push_back(k1); new_children->push_back(k2); }
    auto& par = v_parents_[m];
    if (par[0] < 0) par[0] = e;
    else if (par[1] < 0) par[1] = e;
    else fatal("[AMR] internal error: edge bisected by more than two elements");
}

void AdaptiveMesh::build_vertex_leaves() {
    vleaf_ptr_.assign(x_.size() + 1, 0);
    for (const AmrElem& t : elems_)
        if (t.is_leaf())
            for (int k = 0; k < 3; ++k) vleaf_ptr_[t.v[k] + 1]++;
    for (size_t i = 0; i < x_.size(); ++i) vleaf_ptr_[i + 1] += vleaf_ptr_[i];
    vleaf_list_.resize(vleaf_ptr_[x_.size()]);
    std::vector<int> pos(vleaf_ptr_.begin(), vleaf_ptr_.end() - 1);
    for (int e = 0; e < (int)elems_.size(); ++e)
        if (elems_[e].is_leaf())
            for (int k = 0; k < 3; ++k) vleaf_list_[pos[elems_[e].v[k]]++] = e;
}

// Current leaves containing edge (a, b): descendants (through children containing a) of the
// leaves at a when refinement started, or of the elements bisected through a if a is newer.
void AdaptiveMesh::leaves_with_edge(int a, int b, std::vector<int>& out) {
    std::vector<int>& stack = search_stack_;
    stack.clear();
    if (a + 1 < (int)vleaf_ptr_.size() && vleaf_ptr_[a + 1] > vleaf_ptr_[a]) {
        for (int s = vleaf_ptr_[a]; s < vleaf_ptr_[a + 1]; ++s) stack.push_back(vleaf_list_[s]);
    } else {
        for (int pe : v_parents_[a])
            if (pe >= 0) { stack.push_back(elems_[pe].child[0]); stack.push_back(elems_[pe].child[1]); }
    }
    while (!stack.empty()) {
        const int e = stack.back();
        stack.pop_back();
        if (!has_vertex(e, a)) continue;
        if (elems_[e].is_leaf()) {
            if (has_vertex(e, b)) out.push_back(e);
        } else {
            stack.push_back(elems_[e].child[0]);
            stack.push_back(elems_[e].child[1]);
        }
    }
}

bool AdaptiveMesh::is_hanging(int e) const {
    const AmrElem& t = elems_[e];
    for (int k = 0; k < 3; ++k)
        if (edge_mid_.count(edge_key(t.v[k], t.v[(k + 1) % 3]))) return true;
    return false;
}

// Closure: bisect every hanging leaf (ascending id) near the last bisections until none remain.
int AdaptiveMesh::make_conforming(std::vector<int>& children, std::vector<std::pair<int, int>>& edges) {
    int count = 0;
    std::vector<int> hanging;
    while (true) {
        hanging.clear();
        for (int e : children)
            if (elems_[e].is_leaf()) hanging.push_back(e);
        for (const auto& ab : edges) leaves_with_edge(ab.first, ab.second, hanging);
        std::sort(hanging.begin(), hanging.end());
        hanging.erase(std::unique(hanging.begin(), hanging.end()), hanging.end());
        hanging.erase(std::remove_if(hanging.begin(), hanging.end(), [&](int e) { return !is_hanging(e); }),
                      hanging.end());
        if (hanging.empty()) break;
        children.clear();
        edges.clear();
        for (int e : hanging) {
            bisect(e, &children, &edges);
            ++count;
        }
    }
    return count;
}

// Removes newest vertex m: the children of each element bisected through m are merged back;
// the parents inherit the coarsening flag and the predicted indicators (nodal injection is the
// exact transfer to the coarser space).
void AdaptiveMesh::collapse_vertex(int m, std::vector<char>& coarsen, std::vector<double>& av,
                                   std::vector<double>& ai) {
    for (int s = 0; s < 2; ++s) {
        const int pe = v_parents_[m][s];
        if (pe < 0) continue;
        const int c0 = elems_[pe].child[0], c1 = elems_[pe].child[1];
        const double merged_v = std::sqrt(2.0) * std::max(av[c0], av[c1]);
        const double merged_i = 2.0 * std::max(ai[c0], ai[c1]);
        free_elem(c0);
        free_elem(c1);
        elems_[pe].child[0] = elems_[pe].child[1] = -1;
        coarsen[pe] = 1;
        av[pe] = merged_v;
        ai[pe] = merged_i;
    }
    edge_mid_.erase(v_edge_[m]);
    v_alive_[m] = 0;
    v_parents_[m] = {-1, -1};
    free_verts_.push_back(m);
}

}  // namespace cfa
