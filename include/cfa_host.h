/*
 * C entry points of the host layer (libcfa_host.so): the command-line program run in-process,
 * the figures and per-step history of the last run, and the host mesh / NVB / AMG set-up
 * records for FFI callers (Python ctypes, tests, benchmarks).
 *
 * The reference reaches the same code directly from its main() and its tests
 * (src/cylinder_fem_amr.cu:119-819, tests/test_cfa.cu); these functions are the C-ABI form of
 * those call sites. None of them is on the per-iteration path.
 */
#ifndef CFA_HOST_H
#define CFA_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the command-line program in-process (src/cylinder_fem_amr.cu:119 main); argv[0] is the
 * program name, the flags are those of src/cfa/options.cuh:54-289 plus --mom-tol, --ppe-tol,
 * --proj-tol, --precision fp64|mixed, --quiet, --no-csv. Returns the exit status. */
int cfa_main(int argc, char** argv);

/* out[16] = wall s, momentum s, pressure s, AMR s, AMG set-up s, steps, nodes, elements,
 * non-converged solves, sum of momentum u / v and pressure iterations, last C_D, C_L, delta_p, 0
 * (the figures of the [Timing] lines, src/cylinder_fem_amr.cu:729-783). */
void cfa_last_run_stats(double* out);

/* Per-step history of the last run at full precision: rows of 15 values = step, time, C_D, C_L,
 * delta_p, nodes, elements, p_front, p_back, C_D_surface, C_L_surface (the columns of the
 * forces CSV, src/cylinder_fem_amr.cu:236,526-528), it_u, it_v, it_p, seconds since the time
 * loop started at the end of the step. Copies at most max_rows rows (out may be NULL) and
 * returns the number of rows recorded. */
int cfa_last_run_history(double* out, int max_rows);

/* Kernel timings of the last run with --kernel-timing R (or of the last solver workload): the
 * main kernels launched alone R times each on the final mesh and state, timed with HIP events on
 * the solver stream.
 * Names joined by '\n' into names[cap]; rows of (algorithmic bytes per launch, mean
 * microseconds per launch, launches timed) into out (at most max_rows rows). Returns the number
 * of kernels timed. */
int cfa_last_run_kernels(char* names, int cap, double* out, int max_rows);

/* MMS figures of the last --mms run (src/cylinder_fem_amr.cu:731-783): out[9] = ||u-u_h||_L2,
 * |u-u_h|_H1, ||p-p_h||_L2, ||u||_L2, |u|_H1, ||p||_L2, max |p-p_h| on the cylinder, at the
 * probes, h_max. */
void cfa_last_run_mms(double* out);

/* Base mesh (src/cfa/mesh.cuh:83-193 with n_quarter 8, n_radial 7, n_wake_x 28) refined
 * uniformly `levels` times by NVB (src/cfa/amr.cuh:461-471 refine_uniform), extracted in
 * Morton order (amr.cuh:104-228) and finalised (CSR graph, cylinder edges, validation,
 * src/cfa/mesh.cuh:198-376). Returns an opaque handle for the functions below. */
void* cfa_mesh_uniform(int levels);
void cfa_mesh_free(void* mesh);
/* sizes[6] = nodes, elements, nnz, cylinder edges, max level, (int)(1e6 * min angle in degrees) */
void cfa_mesh_sizes(void* mesh, int* sizes);
/* Copies the arrays; any pointer may be NULL. x, y [n]; bc [n]; ev [3 ne]; rp [n + 1]; ci [nnz];
 * cyl [3 n_cyl] = (a, b, owning element) per cylinder edge. */
void cfa_mesh_arrays(void* mesh, double* x, double* y, int* bc, int* ev, int* rp, int* ci, int* cyl);
/* Owner-computes node -> element table: inc_ptr [n + 1], inc_elem [3 ne], inc_pos [9 ne]. */
void cfa_mesh_incidence(void* mesh, int* inc_ptr, int* inc_elem, int* inc_pos);
/* Cylinder pressure probes at theta = pi and 0 (src/cfa/diagnostics.cuh:23-45):
 * ab = {a_front, b_front, a_back, b_back}, w = {w_front, w_back}. */
void cfa_mesh_probes(void* mesh, int* ab, double* w);

/* Isolated solver workloads on a frozen uniform mesh (the reference's solvers at a chosen size;
 * its momentum study harness scripts/momentum_study does the same on dumped systems).
 * Pressure: the PPE Laplacian (src/cfa/fem_kernels.cuh:201-282 with stab = 0, :460-531 BCs), one
 * SA-AMG set-up, nrhs seeded zero-mean right-hand sides (mt19937_64(1000 + k), N(0,1)) solved
 * from 0 to tol. out[16] = n, nnz, ne, mesh s, AMG set-up ms (a second set-up of the same matrix:
 * buffers sized, kernels loaded, as in a time-stepping run), AMG levels, solves, solve ms,
 * iterations, ms per solve, us per iteration, max final relative residual, first set-up ms, then
 * the same solves with the FP32 cycle: ms per solve, iterations, max final relative residual.
 * Momentum: the BDF2 momentum matrix (fem_kernels.cuh:40-181, :411-453) for a seeded smooth
 * divergence-free field, BiCGSTAB2 with FP64 and with FP32 ILU(0) factors to tol. out[16] = n,
 * nnz, ne, mesh s, colouring ms, colours, FP64 factor ms, solve ms, it_u, it_v, FP32 factor ms,
 * solve ms, it_u, it_v, max |x32 - x64| / max |x64|, us per FP64 iteration.
 * kernel_reps > 0 also times the main kernels (read with cfa_last_run_kernels). */
int cfa_workload_pressure(int level, int nrhs, double tol, int kernel_reps, double* out);
int cfa_workload_momentum(int level, double tol, int kernel_reps, double* out);
/* AMG set-up of the last pressure workload (its second, timed set-up): out[12] = ms in download, D^-1 + spectral radius,
 * aggregation, P smoothing, transpose, Galerkin product, upload, coarse inverse, remap, then the
 * number of levels built on the device, the rounds of the aggregation's decision front and the
 * sparse products that took the global-memory sort. */
int cfa_workload_amg_phases(double* out);
/* Host copies of the workload systems, for checks outside the library (no counterpart in the
 * reference, which has no isolated workloads). cfa_workload_keep_systems(1): the following
 * workload calls keep their matrix, right-hand sides, initial guesses and solutions (0 drops
 * them). cfa_workload_system_sizes: n, nnz and right-hand sides of system `which` (0 the
 * pressure workload, 1 the momentum workload). cfa_workload_system copies rp[n+1], ci[nnz],
 * tag[n], val[nnz], b[cols n], x0[cols n] and x: pressure [cols n], the solution of each
 * right-hand side; momentum [4 n], u and v with FP64 factors, then with FP32 factors; and the
 * Krylov iterations of each of these solutions in iters. */
int cfa_workload_keep_systems(int on);
int cfa_workload_system_sizes(int which, int* n, int* nnz, int* cols);
int cfa_workload_system(int which, int* rp, int* ci, int* tag, double* val, double* b, double* x0, double* x,
                        int* iters);

/* Solver-study dumps written by --dump-systems N (src/cylinder_fem_amr.cu:331-352 and :375-378:
 * <prefix>_momentum.bin = int32 n, nnz; row_ptr [n + 1], col_ind [nnz] int32; values [nnz],
 * rhs_u, rhs_v, u0, v0, x, y [n] float64; bc_tag [n] int32; then the solutions u, v [n];
 * :426-438: <prefix>_pressure.bin = int32 n, nnz; row_ptr, col_ind; K_p values, rhs, initial
 * guess). The reference reads them in scripts/momentum_study (study.py load(), mb.cu, mb_amg.cu).
 * All return CFA_E_IO for a missing file or one whose size or CSR pattern does not fit.
 * cfa_dump_header: n and nnz of either file.
 * cfa_dumped_momentum_solve: BiCGSTAB2 with multicolour ILU(0) (precision 0: FP64 factors,
 * 1: FP32 factors) from the dumped u0, v0 to tol; x_u, x_v [n] (host, may be NULL) receive the
 * solution. out[11] = n, nnz, colours, colouring ms, factor ms, solve ms, it_u, it_v,
 * converged_u, converged_v, max relative deviation from the dumped solution (-1 if the file has
 * none).
 * cfa_dumped_pressure_solve: PCG with SA-AMG (precision 1: FP32 cycle, FP64 coarse solve) on K_p
 * alone (the files hold no OSS operator) from the dumped guess, or from 0 with from_zero = 1 (as
 * mb_amg.cu); x [n] may be NULL. out[8] = n, nnz, AMG levels, set-up ms, solve ms, iterations,
 * converged, final relative residual.
 * kernel_reps > 0 also times the SpMV, the ILU apply or the V-cycle (cfa_last_run_kernels). */
int cfa_dump_header(const char* path, int* n, int* nnz);
int cfa_dumped_momentum_solve(const char* path, double tol, int max_iter, int precision, int kernel_reps,
                              double* x_u, double* x_v, double* out);
int cfa_dumped_pressure_solve(const char* path, double tol, int max_iter, int precision, int from_zero,
                              int kernel_reps, double* x, double* out);

/* DFG 2D-2 statistics of a force history, as the program reports them at the end of a run
 * (src/cfa/diagnostics.cuh:47-121 ForceHistory::compute; window = --stats-window):
 * out[13] = number of lift periods, the 6 period means (C_D max, C_D min, C_L max, C_L min,
 * Strouhal number, delta_p half a period after the lift maximum), then their 6 spreads. */
int cfa_force_statistics(int n, const double* t, const double* cd, const double* cl, const double* dp,
                         double window, double* out);

/* Host SA-AMG set-up (src/cfa/linalg.cuh:995-1060, 1276-1431) of a host CSR matrix: per level
 * rows, (nnz A, nnz P), Jacobi weight, the coarse size and the level-0 aggregates [n].
 * Returns the number of smoothed levels (<= 16 reported). Needs a GPU (the hierarchy is
 * uploaded as in the solver). */
int cfa_amg_setup_record(int n, const int* rp, const int* ci, const double* val, int* level_rows,
                         int* level_nnz, double* level_omega, int* coarse_n, int* agg0);

/* Discrete adjoint of the projection step on a frozen uniform mesh (no counterpart in the
 * reference, which has no derivative code). The step is the program's time step
 * (src/cylinder_fem_amr.cu:277-538, default options: BDF2 after step 1, rotational projection,
 * OSS 0.3, SUPG time scale 1e-3, ILU(0)-BiCGSTAB2 momentum, SA-AMG PCG pressure rebuilt every 50
 * steps, do-nothing outlet) with parameters theta = {viscosity nu, inflow peak velocity U_max}
 * (the reference's constants are nu = 1e-3, U_max = 1.5) and both solvers run to relative
 * tolerance tol. A state is 6 n doubles [u | v | u_old | v_old | p | dp] (host arrays; dp is the
 * pressure solver's warm start). `step` numbers are the program's (step 1 is the BDF1 start).
 * The window functional is J = (1/W) sum over W steps of (weights[0] C_D + weights[1] C_L), the
 * volume-integral force coefficients (src/cfa/fem_kernels.cuh:1046-1123).
 * cfa_adjoint_create: level 0..6 uniform NVB refinements of the base mesh; NULL on failure.
 * cfa_adjoint_initial_state: the DFG 2D-2 initial condition (src/cfa/device.cuh:466-486).
 * cfa_step_fwd: one step, forces[2] = (C_D, C_L) of the new state (may be NULL).
 * cfa_window_forward: W steps from state0; state_out (may be NULL), J, history [2 W] (may be
 * NULL) = (C_D, C_L) per step.
 * cfa_adjoint_set_shape: deform the hole boundary by 64 radial modes, r(phi) = R (1 + sum_k
 * eps_k g_k(phi)), g_k Gaussians centred at 2 pi k / 64 of width 2 pi / 64, carried into the
 * mesh by a smooth radial cut-off that vanishes at 3 R (n_modes must be 64; all zero = the
 * reference's geometry). The deformation holds for the following calls.
 * cfa_window_adjoint: J and dJ/dtheta [2] by the discrete adjoint with checkpoints every
 * checkpoint_every steps (0: one segment) and recomputation in between; dJ_dshape [64] (may be
 * NULL) = dJ/deps at the current shape; dJ_dstate0 [6 n] (may be NULL) = dJ/dstate0 (zero in the
 * dp slot); stats [5] (may be NULL) = forward ms, recompute + reverse ms, recomputed steps,
 * adjoint PCG iterations, adjoint BiCGSTAB2 iterations. */
typedef struct cfa_adjoint cfa_adjoint_t;
cfa_adjoint_t* cfa_adjoint_create(int level, double dt, double tol);
void cfa_adjoint_destroy(cfa_adjoint_t* ctx);
int cfa_adjoint_sizes(const cfa_adjoint_t* ctx, int* n, int* ne, int* nnz);
int cfa_adjoint_initial_state(const cfa_adjoint_t* ctx, double* state);
int cfa_step_fwd(cfa_adjoint_t* ctx, const double* theta, int step, const double* state_in, double* state_out,
                 double* forces);
int cfa_window_forward(cfa_adjoint_t* ctx, const double* theta, int first_step, int steps, const double* weights,
                       const double* state0, double* state_out, double* J, double* history);
int cfa_adjoint_set_shape(cfa_adjoint_t* ctx, const double* eps, int n_modes);
int cfa_window_adjoint(cfa_adjoint_t* ctx, const double* theta, int first_step, int steps, int checkpoint_every,
                       const double* weights, const double* state0, double* J, double* dJ_dtheta,
                       double* dJ_dshape, double* dJ_dstate0, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* CFA_HOST_H */
