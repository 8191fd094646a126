// ORACLE (test infrastructure only): C entry points for the Python tests (ctypes). Arrays are
// caller-owned host buffers; each function restates one reference routine (see the C++ side).
#include <cmath>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cfa_oracle.hpp"

using namespace oracle;

extern "C" {

// Fixed-mesh run. rows: steps x 13 doubles (step, t, C_D, C_L, dP, nodes, elements, p_front,
// p_back, C_D_surface, C_L_surface, it_u + 1000 it_v, it_p); mms: 9 doubles (6 norms, cyl, probe,
// h_max). With a prefix the forces/MMS CSV files are written, and with dump_systems > 0 the
// systems of that step (<prefix>_momentum.bin, <prefix>_pressure.bin). Returns the number of rows.
int oracle_run_fixed(int steps, double dt, int uniform_level, int min_level, int mms, double mom_tol,
                     double ppe_tol, const char* prefix, int dump_systems, double* rows, double* mms_out,
                     double* wall_s) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = min_level;
    c.mms = mms != 0;
    c.mom_tol = mom_tol;
    c.ppe_tol = ppe_tol;
    c.quiet = true;
    if (prefix) c.prefix = prefix;
    c.dump_systems = dump_systems;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    if (mms_out) {
        for (int k = 0; k < 6; ++k) mms_out[k] = r.mms[k];
        mms_out[6] = r.cyl_max;  mms_out[7] = r.probe_max;  mms_out[8] = r.h_max;
    }
    if (wall_s) *wall_s = r.wall_s;
    return (int)r.rows.size();
}

// Fixed-mesh DFG flow (no MMS) from the initial condition with the step's viscosity and inflow
// peak as parameters, both solvers at relative tolerance tol: rows as oracle_run_fixed.
int oracle_run_params(int steps, double dt, int uniform_level, double nu, double u_max, double tol, double* rows) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = 0;
    c.nu = nu;
    c.u_max = u_max;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    return (int)r.rows.size();
}

// Fixed-mesh DFG run with the outlet parabola (--outlet dirichlet), both solvers at relative
// tolerance tol, case 2D-2 or (case_2d3) 2D-3: rows as oracle_run_fixed.
int oracle_run_outlet_dirichlet(int steps, double dt, int uniform_level, int min_level, int case_2d3, double tol,
                                double* rows) {
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.uniform_level = uniform_level;
    c.min_level = min_level;
    c.case_2d3 = case_2d3 != 0;
    c.outlet_dirichlet = true;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_fixed(c);
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    return (int)r.rows.size();
}

// Continues a run on a fixed mesh from a --dump-state file for `steps` steps (DFG inflow, 2D-2,
// default flags), both solvers at relative tolerance tol, on `threads` OpenMP threads (<= 0: the
// runtime's default): rows as oracle_run_fixed, wall seconds of the time loop in *wall_s.
int oracle_run_state(const char* path, int steps, double dt, double tol, int threads, double* rows, double* wall_s) {
#ifdef _OPENMP
    const int prev = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#endif
    RunConfig c;
    c.steps = steps;
    c.dt = dt;
    c.mom_tol = tol;
    c.ppe_tol = tol;
    c.quiet = true;
    RunResult r = run_state_file(c, path);
#ifdef _OPENMP
    omp_set_num_threads(prev);
#endif
    for (size_t k = 0; k < r.rows.size(); ++k) {
        const StepRow& s = r.rows[k];
        double* o = rows + 13 * k;
        o[0] = s.step;  o[1] = s.t;  o[2] = s.cd;  o[3] = s.cl;  o[4] = s.dp;  o[5] = s.nodes;
        o[6] = s.elems;  o[7] = s.p_front;  o[8] = s.p_back;  o[9] = s.cd_s;  o[10] = s.cl_s;
        o[11] = s.it_u + 1000.0 * s.it_v;  o[12] = s.it_p;
    }
    if (wall_s) *wall_s = r.wall_s;
    return (int)r.rows.size();
}

// ---- standalone solves of a given system (the solver workloads and the Jacobi paths) ----------
namespace {
Csr make_csr(int n, const int* rp, const int* ci, const double* val) {
    Csr A;
    A.rows = A.cols = n;
    A.rp.assign(rp, rp + n + 1);
    A.ci.assign(ci, ci + rp[n]);
    A.v.assign(val, val + rp[n]);
    return A;
}
}  // namespace

// SA-AMG set-up on A, then PCG (SA-AMG V-cycle, no extra operator) from x (in/out) to relative
// residual tol (src/cfa/linalg.cuh:172-256, :864-1432). stats[4] = iterations, converged, final
// relative residual, AMG levels. Returns the iteration count.
int oracle_pcg_amg(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, double tol,
                   int max_iter, double* stats) {
    Amg amg;
    amg.setup(make_csr(n, rp, ci, val));
    Op prec = [&](const double* r, double* z) { amg.apply(r, z); };
    const SolveInfo si = pcg(n, rp, ci, val, b, x, max_iter, tol, nullptr, &prec);
    if (stats) {
        stats[0] = si.iters;  stats[1] = si.converged ? 1.0 : 0.0;  stats[2] = si.rel;  stats[3] = amg.num_levels();
    }
    return si.iters;
}

// PCG with the Jacobi preconditioner z = r / diag(A) (--pressure-precond jacobi,
// src/cfa/linalg.cuh:172-256 with the inverse diagonal of :211). stats as oracle_pcg_amg.
int oracle_pcg_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b, double* x, double tol,
                      int max_iter, double* stats) {
    Vec dinv(n, 1.0);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { dinv[i] = std::fabs(val[k]) > 1.0e-14 ? 1.0 / val[k] : 1.0; break; }
    Op prec = [&](const double* r, double* z) { for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i]; };
    const SolveInfo si = pcg(n, rp, ci, val, b, x, max_iter, tol, nullptr, &prec);
    if (stats) {
        stats[0] = si.iters;  stats[1] = si.converged ? 1.0 : 0.0;  stats[2] = si.rel;  stats[3] = 1.0;
    }
    return si.iters;
}

// Greedy multicolour ILU(0) of A and the two-component BiCGSTAB2 from (xu, xv) (in/out) to
// relative residual tol (src/cfa/linalg.cuh:361-779). stats[7] = iterations u, v, converged u, v,
// final relative residuals u, v, colours. Returns 0.
int oracle_bicgstab2_ilu(int n, const int* rp, const int* ci, const double* val, const double* bu, const double* bv,
                         double* xu, double* xv, double tol, int max_iter, double* stats) {
    Ilu0 ilu;
    ilu.set_pattern(n, rp, ci);
    ilu.factor(val);
    const Bicg2Info bi = bicgstab2(n, rp, ci, val, bu, bv, xu, xv, max_iter, tol, ilu);
    if (stats) {
        stats[0] = bi.iters[0];  stats[1] = bi.iters[1];
        stats[2] = bi.converged[0] ? 1.0 : 0.0;  stats[3] = bi.converged[1] ? 1.0 : 0.0;
        stats[4] = bi.rel[0];  stats[5] = bi.rel[1];  stats[6] = ilu.n_colours;
    }
    return 0;
}

// Jacobi-preconditioned BiCGSTAB from x (in/out) (--momentum-precond jacobi,
// src/cfa/linalg.cuh:261-347). stats[3] = iterations, converged, final relative residual.
int oracle_bicgstab_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
                           double tol, int max_iter, double* stats) {
    const SolveInfo si = bicgstab_jacobi(n, rp, ci, val, b, x, max_iter, tol);
    if (stats) {
        stats[0] = si.iters;  stats[1] = si.converged ? 1.0 : 0.0;  stats[2] = si.rel;
    }
    return si.iters;
}

// The base mesh (src/cylinder_fem_amr.cu:169) refined uniformly `levels` times by the oracle's NVB
// restatement, extracted in Morton order and finalised (CSR graph, cylinder edges, checks); an
// opaque handle for oracle_mesh_sizes / oracle_mesh_arrays / oracle_mesh_free.
void* oracle_mesh_uniform(int levels) {
    Mesh2D* m = new Mesh2D();
    cfa::generate_cylinder_mesh(*m, 8, 7, 28);
    cfa::finalize_mesh(*m);
    Vec z(m->num_nodes, 0.0), u, v, p, uo, vo;
    NvbForest amr(*m, z, z, z);
    amr.extract(*m, u, v, p, uo, vo);
    cfa::finalize_mesh(*m);
    if (levels > 0) {
        amr.refine_uniform(levels);
        amr.extract(*m, u, v, p, uo, vo);
        cfa::finalize_mesh(*m);
    }
    return m;
}

void oracle_mesh_free(void* h) { delete static_cast<Mesh2D*>(h); }

// sizes[6] = nodes, elements, nnz, cylinder edges, max level, (int)(1e6 * min angle deg)
void oracle_mesh_sizes(void* h, int* sizes) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    cfa::MeshQuality q;
    cfa::validate_mesh(m, q);
    sizes[0] = m.num_nodes;
    sizes[1] = m.num_elems;
    sizes[2] = m.nnz;
    sizes[3] = (int)m.cylinder_edges.size();
    sizes[4] = q.max_level;
    sizes[5] = (int)(q.min_angle_deg * 1.0e6);
}

// x, y [n], bc [n], ev [3 ne], rp [n + 1], ci [nnz], cyl [3 n_cyl] (a, b, element)
void oracle_mesh_arrays(void* h, double* x, double* y, int* bc, int* ev, int* rp, int* ci, int* cyl) {
    const Mesh2D& m = *static_cast<Mesh2D*>(h);
    std::memcpy(x, m.x.data(), m.x.size() * sizeof(double));
    std::memcpy(y, m.y.data(), m.y.size() * sizeof(double));
    std::memcpy(bc, m.bc_tag.data(), m.bc_tag.size() * sizeof(int));
    for (int e = 0; e < m.num_elems; ++e)
        for (int k = 0; k < 3; ++k) ev[3 * e + k] = m.elems[e].v[k];
    std::memcpy(rp, m.csr_row_ptr.data(), m.csr_row_ptr.size() * sizeof(int));
    std::memcpy(ci, m.csr_col_ind.data(), m.csr_col_ind.size() * sizeof(int));
    for (size_t k = 0; k < m.cylinder_edges.size(); ++k) {
        cyl[3 * k] = m.cylinder_edges[k].first;
        cyl[3 * k + 1] = m.cylinder_edges[k].second;
        cyl[3 * k + 2] = m.cylinder_edge_elem[k];
    }
}

}  // extern "C"
