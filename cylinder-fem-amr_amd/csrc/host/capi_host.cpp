// C entry points of the host layer for FFI callers (Python ctypes, tests): meshes of the
// reference's generator and NVB refinement, their CSR graph and the owner-computes incidence
// table, and the host AMG set-up / colouring records.
#include <cstring>

#include "device.hpp"
#include "outputs.hpp"

using namespace cfa;

extern "C" {

// Base mesh (src/cylinder_fem_amr.cu:169) refined uniformly `levels` times by NVB, extracted in
// Morton order and finalised (CSR, cylinder edges, validation). Returns an opaque handle.
void* cfa_mesh_uniform(int levels) {
    Mesh2D* m = new Mesh2D();
    generate_cylinder_mesh(*m, 8, 7, 28);
    finalize_mesh(*m);
    std::vector<double> u(m->num_nodes, 0.0), v = u, p = u, uo = u, vo = u;
    AdaptiveMesh amr(*m, u, v, p);
    amr.extract(*m, u, v, p, uo, vo);
    finalize_mesh(*m);
    if (levels > 0) {
        amr.refine_uniform(levels);
        amr.extract(*m, u, v, p, uo, vo);
        finalize_mesh(*m);
    }
    return m;
}

void cfa_mesh_free(void* h) { delete static_cast<Mesh2D*>(h); }

// sizes[6] = nodes, elements, nnz, cylinder edges, max level, (int)(1e6 * min angle deg)
void cfa_mesh_sizes(void* h, int* sizes) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    MeshQuality q;
    validate_mesh(m, q);
    sizes[0] = m.num_nodes;
    sizes[1] = m.num_elems;
    sizes[2] = m.nnz;
    sizes[3] = (int)m.cylinder_edges.size();
    sizes[4] = q.max_level;
    sizes[5] = (int)(q.min_angle_deg * 1.0e6);
}

// Copies the arrays; any pointer may be NULL. cyl: [3 n_cyl] (a, b, element).
void cfa_mesh_arrays(void* h, double* x, double* y, int* bc, int* ev, int* rp, int* ci, int* cyl) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    if (x) std::memcpy(x, m.x.data(), m.x.size() * sizeof(double));
    if (y) std::memcpy(y, m.y.data(), m.y.size() * sizeof(double));
    if (bc) std::memcpy(bc, m.bc_tag.data(), m.bc_tag.size() * sizeof(int));
    if (ev)
        for (int e = 0; e < m.num_elems; ++e)
            for (int k = 0; k < 3; ++k) ev[3 * e + k] = m.elems[e].v[k];
    if (rp) std::memcpy(rp, m.csr_row_ptr.data(), m.csr_row_ptr.size() * sizeof(int));
    if (ci) std::memcpy(ci, m.csr_col_ind.data(), m.csr_col_ind.size() * sizeof(int));
    if (cyl)
        for (size_t k = 0; k < m.cylinder_edges.size(); ++k) {
            cyl[3 * k] = m.cylinder_edges[k].first;
            cyl[3 * k + 1] = m.cylinder_edges[k].second;
            cyl[3 * k + 2] = m.cylinder_edge_elem[k];
        }
}

// Owner-computes incidence: inc_ptr [n + 1], inc_elem [3 ne], inc_pos [9 ne].
void cfa_mesh_incidence(void* h, int* inc_ptr, int* inc_elem, int* inc_pos) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    std::vector<int> p, e, q;
    build_incidence(m, p, e, q);
    std::memcpy(inc_ptr, p.data(), p.size() * sizeof(int));
    std::memcpy(inc_elem, e.data(), e.size() * sizeof(int));
    std::memcpy(inc_pos, q.data(), q.size() * sizeof(int));
}

// Pressure probes at theta = pi (front) and 0 (back): out_ab = {a0, b0, a1, b1}, out_w = {w0, w1}
void cfa_mesh_probes(void* h, int* out_ab, double* out_w) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    const PressureProbe f = find_cylinder_probe(m, Sim::PI), b = find_cylinder_probe(m, 0.0);
    out_ab[0] = f.a; out_ab[1] = f.b; out_ab[2] = b.a; out_ab[3] = b.b;
    out_w[0] = f.w; out_w[1] = b.w;
}

// Host SA-AMG set-up records of a CSR matrix (host arrays): level sizes (rows, nnz of A, P) for
// up to 16 levels, the coarse size, the Jacobi weights and the level-0 aggregates.
// Returns the number of smoothed levels. Needs an initialised device (the hierarchy is uploaded).
int cfa_amg_setup_record(int n, const int* rp, const int* ci, const double* val, int* level_rows,
                         int* level_nnz, double* level_omega, int* coarse_n, int* agg0) {
    DevArray<int> d_rp, d_ci;
    DevArray<double> d_v;
    d_rp.upload(rp, n + 1);
    d_ci.upload(ci, rp[n]);
    d_v.upload(val, rp[n]);
    AmgPreconditioner amg;
    amg.keep_aggregates = agg0 != nullptr;
    amg.setup(n, rp[n], d_rp, d_ci, d_v);
    const int nl = (int)amg.levels.size();
    for (int l = 0; l < nl && l < 16; ++l) {
        level_rows[l] = amg.levels[l]->A.rows;
        level_nnz[2 * l] = amg.levels[l]->A.nnz;
        level_nnz[2 * l + 1] = amg.levels[l]->P.nnz;
        level_omega[l] = amg.levels[l]->omega;
    }
    *coarse_n = amg.coarse_n;
    if (agg0 && !amg.last_aggregates.empty())
        std::memcpy(agg0, amg.last_aggregates.data(), amg.last_aggregates.size() * sizeof(int));
    return nl;
}

// DFG 2D-2 statistics of a force history (the program's end-of-run report):
// out[13] = lift periods, 6 mean values, 6 spreads (C_D max/min, C_L max/min, St, dp half period).
int cfa_force_statistics(int n, const double* t, const double* cd, const double* cl, const double* dp, double window,
                         double* out) {
    if (n < 0 || !out || (n > 0 && (!t || !cd || !cl || !dp))) return CFA_E_ARG;
    ForceHistory h;
    for (int i = 0; i < n; ++i) h.add(t[i], cd[i], cl[i], dp[i]);
    size_t first = 0;
    const ForceHistory::Stats S = h.statistics(window, first);
    out[0] = S.periods;
    for (int k = 0; k < 6; ++k) {
        out[1 + k] = S.value[k];
        out[7 + k] = S.spread[k];
    }
    return CFA_OK;
}

}  // extern "C"
