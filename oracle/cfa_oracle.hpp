// ORACLE — test infrastructure only. A serial FP64 CPU restatement of the reference's per-step
// projection solve, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker. The product (cylinder-fem-amr_amd/) never includes, links or runs anything here.
//
// What it restates (each function cites the reference code it follows):
//   * element kernels of src/cfa/fem_kernels.cuh (assembly, BCs, OSS, correction, forces, MMS,
//     indicators) as element loops with plain scatter-add, elements in mesh order;
//   * the solvers of src/cfa/linalg.cuh (PCG, BiCGSTAB, BiCGSTAB2, multicolour ILU(0), SA-AMG
//     setup and V-cycle) with the same stopping rules, guards and constants; dot products are
//     plain left-to-right sums;
//   * the time loop of src/cylinder_fem_amr.cu:277-538 and the re-projection of
//     src/cfa/device.cuh:337-464.
// Meshes: the base mesh from the product's generator (pinned by the reference's stored mesh sizes),
// refined by the oracle's own restatement of the NVB rules (cfa_oracle_amr.hpp).
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "cfa_oracle_amr.hpp"

namespace oracle {

using cfa::Mesh2D;
using Vec = std::vector<double>;

struct Csr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    Vec v;
};

// ---- element kernels (src/cfa/fem_kernels.cuh) --------------------------------------------
// Outlet condition (src/cfa/common.cuh:76-89): default p = 0 on outlet nodes; with dirichlet the
// outlet carries the inflow parabola times scale and dp = 0 at node pin only.
struct Outlet {
    bool dirichlet = false;
    int pin = -1;
    double scale = 1.0;
    bool p_dirichlet(int i, int tag) const { return dirichlet ? i == pin : tag == cfa::BC_OUTLET; }
};
// src/cfa/device.cuh:153-172: the pinned node and the outlet parabola factor of a mesh
Outlet outlet_dirichlet_setup(const Mesh2D& mesh);
// src/cfa/device.cuh:179-199: a transferred field with its outlet values flux-balanced
Vec flux_balanced(const Mesh2D& mesh, const Vec& f);
// src/cfa/fem_kernels.cuh:493-510 k_sum + k_subtract_mean: f -= (sum f / sum m) m, each sum by
// blocks of 256 (in-block tree), the block results added in ascending order
void subtract_mean(int n, double* f, const double* m);

struct MmsValues { double u, v, p, fx, fy, ux, uy, vx, vy; };
MmsValues mms_exact(double x, double y, double t);

int find_csr(const int* ci, int s, int e, int col);

void assemble_momentum(int ne, const int* ev, const double* x, const double* y,
                       const double* u_old, const double* v_old, const double* u_prev,
                       const double* v_prev, const double* p_old, double nu, double rho,
                       double a0, double a1, double a2, double c1, double c2, double tau_time,
                       double skew, const int* rp, const int* ci, double* A, double* rhs_u,
                       double* rhs_v, double* mass, const double* fx, const double* fy);
void apply_bc_momentum(int n, const int* tag, const double* x, const double* y, double u_max,
                       double mms_time, const int* rp, const int* ci, double* A, double* rhs_u,
                       double* rhs_v, double* u, double* v, const Outlet& outlet = Outlet());
void mms_forcing(int n, const double* x, const double* y, double t, double* fx, double* fy);
void compute_oss_weights(int ne, const int* ev, const double* x, const double* y, const double* us,
                         const double* vs, double dt, double rho, double nu, double stab, double* s);
void accumulate_scalar_gradient(int ne, const int* ev, const double* x, const double* y,
                                const double* f, const double* s, double* gx, double* gy);
void assemble_pressure_poisson(int ne, const int* ev, const double* x, const double* y,
                               const double* us, const double* vs, const double* p,
                               const double* gpx, const double* gpy, const double* mass,
                               const double* s, double dt, double rho, double stab, const int* rp,
                               const int* ci, double* K, double* rhs, double* div_num);
void apply_bc_pressure(int n, const int* tag, const int* rp, const int* ci, double* K, double* rhs,
                       double* dp, const Outlet& outlet = Outlet());
void assemble_lumped_mass(int ne, const int* ev, const double* x, const double* y, double* m);
void assemble_oss_operator(int ne, const int* ev, const double* x, const double* y, const double* s,
                           const int* rp, const int* ci, int nnz, double* oss_val);
void oss_apply(int n, const int* rp, const int* ci, int nnz, const double* oss_val,
               const double* inv_m, const int* tag, const double* in, double* out,
               const Outlet& outlet = Outlet());
void oss_apply_matrix_free(int ne, int n, const int* ev, const double* x, const double* y,
                           const int* tag, const double* s, const double* mass, const double* in,
                           double* out, const Outlet& outlet = Outlet());
void accumulate_grad_dp_and_vorticity(int ne, const int* ev, const double* x, const double* y,
                                      const double* dp, const double* u, const double* v,
                                      double* gx, double* gy, double* vort_num);
void velocity_pressure_correction(int n, const int* tag, double dt, double rho, double p_update,
                                  double rot, const double* div_num, const double* mass,
                                  const double* gx, const double* gy, const double* vort_num,
                                  const double* dp, double* u, double* v, double* p, double* vort,
                                  const Outlet& outlet = Outlet());
void stiffness_apply(int ne, const int* ev, const double* x, const double* y, const double* f,
                     double c, double* out);
void volume_forces(int ne, const int* ev, const int* tag, const double* x, const double* y,
                   const double* u, const double* v, const double* u_old, const double* v_old,
                   const double* u_prev, const double* v_prev, const double* p, double b0,
                   double b1, double b2, double nu, double rho, double* out);
void surface_forces(int nedges, const int* edges, const int* ev, const double* x, const double* y,
                    const double* u, const double* v, const double* p, double nu, double rho,
                    double xc, double yc, double* out);
void mms_errors(int ne, const int* ev, const double* x, const double* y, const double* uh,
                const double* vh, const double* ph, double t, double* out);
void error_indicator(int ne, int n, const int* ev, const double* x, const double* y,
                     const double* u, const double* v, double u_max, double* mass,
                     double* grad_rec, double* elem_error);

// ---- solvers (src/cfa/linalg.cuh) -----------------------------------------------------------
double dot(int n, const double* a, const double* b);
void spmv(int n, const int* rp, const int* ci, const double* val, const double* x, double* y);

using Op = std::function<void(const double*, double*)>;
struct SolveInfo { int iters = 0; bool converged = true; double rel = 0.0; };
SolveInfo pcg(int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
              int max_iter, double tol, const Op* extra, const Op* prec);
SolveInfo bicgstab_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b,
                          double* x, int max_iter, double tol);

struct Ilu0 {
    int n = 0, n_colours = 0;
    std::vector<int> colour, colour_start, rows, diag;
    std::vector<int> lp, up, lc, uc, l_src, u_src, d_src;
    Vec lu, lv, uv, dinv;
    const int *rp = nullptr, *ci = nullptr;
    void set_pattern(int n, const int* rp, const int* ci);
    void factor(const double* val);
    void apply(const double* r, double* z, int nrhs) const;
};
struct Bicg2Info { int iters[2] = {0, 0}; bool converged[2] = {true, true}; double rel[2] = {0, 0}; };
Bicg2Info bicgstab2(int n, const int* rp, const int* ci, const double* val, const double* bu,
                    const double* bv, double* xu, double* xv, int max_iter, double tol,
                    const Ilu0& ilu);

struct Amg {
    int smooth_steps = 2;
    double theta = 0.08;
    int coarse_target = 300;
    int max_levels = 15;
    struct Level { Csr A, P, R; Vec dinv; double omega = 0; Vec x, b, r, Ax; };
    std::vector<Level> levels;
    int coarse_n = 0;
    Vec coarse_inv, cb, cx;
    void setup(const Csr& A0);
    void apply(const double* r, double* z);
    int num_levels() const { return (int)levels.size() + 1; }
private:
    void vcycle(int l, const double* b);
};
double spectral_radius_DinvA(const Csr& A, const Vec& dinv);
Csr csr_multiply(const Csr& A, const Csr& B);
Csr csr_transpose(const Csr& A);
Csr smoothed_prolongator(const Csr& A, const Vec& dinv, double omega, double theta, std::vector<int>* agg_out);
Vec dense_spd_inverse(Vec M, int n);
Csr amg_remap_p0(const Csr& P0, int n, const int* rp, const int* ci, const std::vector<int>& new_to_old, bool* ok);

// ---- driver (src/cylinder_fem_amr.cu) ---------------------------------------------------------
struct RunConfig {
    int steps = 100;
    double dt = 1.0e-3;
    int uniform_level = 0;
    int min_level = 2;
    bool mms = false;
    bool bdf2 = true;
    bool rotational = true;
    bool case_2d3 = false;
    double pressure_stab = 0.3;
    double supg_tau_time = 1.0e-3;
    int amg_rebuild = 50;
    double mom_tol = 1.0e-7;
    double ppe_tol = 1.0e-7;
    double proj_tol = 1.0e-8;
    int print_interval = 10;
    bool quiet = false;
    std::string prefix;   // empty: no files
    int dump_systems = 0;  // > 0 (with a prefix): write the momentum and pressure systems of this step
    double nu = cfa::Sim::NU;        // viscosity of the step (the reference's constant by default)
    double u_max = -1.0;             // > 0: inflow peak of the Dirichlet data instead of inflow_peak(t)
    bool outlet_dirichlet = false;   // --outlet dirichlet
};
struct StepRow {
    int step;
    double t, cd, cl, dp, p_front, p_back, cd_s, cl_s;
    int nodes, elems, it_u, it_v, it_p;
};
struct RunResult {
    std::vector<StepRow> rows;
    double mms[6] = {0, 0, 0, 0, 0, 0};   // sqrt of the k_mms_errors sums
    double cyl_max = 0, probe_max = 0, h_max = 0;
    double wall_s = 0;
    int nonconverged = 0;
};
// Fixed (uniform) mesh run, optionally starting from `initial_mesh_levels` (the product mesh)
RunResult run_fixed(const RunConfig& cfg);
// cfg.steps steps on a fixed mesh from a given state (uⁿ, vⁿ, pⁿ, uⁿ⁻¹, vⁿ⁻¹ and the pressure
// increment as warm start) after step first_step - 1; wall time counted from t_wall0
RunResult run_loop(const RunConfig& cfg, Mesh2D& mesh, Vec hu, Vec hv, Vec hp, Vec huo, Vec hvo, Vec dp0,
                   int first_step, cfa::Clock::time_point t_wall0);
// run_loop from a state file written by the program's --dump-state (mesh and fields of a step)
RunResult run_state_file(const RunConfig& cfg, const std::string& path);

}  // namespace oracle
