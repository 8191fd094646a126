/*
 * C ABI of the MI355X (gfx950) kernels for the per-step projection solve of cylinder_fem_amr.
 *
 * The reference has no plugin ABI: its solver path is a set of C++ classes and free functions
 * inside one CUDA translation unit (src/cylinder_fem_amr.cu includes the eight src/cfa headers). Each entry
 * point below replaces one kernel group of that path; the comment on it names the reference
 * code it stands in for. The C++ host layer (Mesh2D, AdaptiveMesh, GpuLinearSolver,
 * MulticolourIlu0, GpuBicgstab2, AmgPreconditioner, DeviceMeshBuffers, the CLI) calls only
 * these functions, and so can a Python or other FFI caller.
 *
 * Conventions
 *   - All pointers named d_* or documented as "device" are device memory; sizes are element
 *     counts unless they say bytes. Indices are int32, values FP64 unless a name ends in _f32.
 *   - Every call is asynchronous on the given stream (NULL = the default stream) unless noted.
 *   - Return value: 0 on success, a HIP error code (> 0) on a runtime failure, or a negative
 *     CFA_E_* code for an invalid argument. cfa_error_string() describes any code.
 *   - No entry point allocates device memory except cfa_malloc / cfa_host_malloc: scratch
 *     comes from caller-provided workspaces.
 */
#ifndef CFA_HIP_H
#define CFA_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* cfa_stream_t; /* hipStream_t */

enum {
    CFA_OK = 0,
    CFA_E_ARG = -1,      /* invalid argument (null pointer, negative size, ...) */
    CFA_E_ROWLEN = -2,   /* a CSR row longer than the kernel limit (ILU: 32, as linalg.cuh:361) */
    CFA_E_NOTREADY = -3, /* object not set up */
    CFA_E_IO = -4,       /* file missing or not in the expected layout */
    CFA_E_SIZE = -5      /* input larger than the fixed workspace of a no-allocation entry point */
};

/* ---------------------------------------------------------------------------------------------
 * Runtime plumbing (replaces the CUDA runtime calls and CUDA_CHECK of src/cfa/common.cuh:33-43)
 * ------------------------------------------------------------------------------------------- */
int cfa_init(int device);
const char* cfa_error_string(int code);
int cfa_malloc(void** d_ptr, size_t bytes);
int cfa_free(void* d_ptr);
/* pinned host memory, mapped into the device address space (d_alias may be NULL) */
int cfa_host_malloc(void** h_ptr, void** d_alias, size_t bytes);
int cfa_host_free(void* h_ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int cfa_memcpy(void* dst, const void* src, size_t bytes, int kind, cfa_stream_t stream);
/* zero fill by a kernel (no memset node, safe inside captured graphs) */
int cfa_zero(void* d_ptr, size_t bytes, cfa_stream_t stream);
int cfa_stream_sync(cfa_stream_t stream);
/* a stream that does not synchronise with the default stream (default priority) */
int cfa_stream_create(cfa_stream_t* stream);
int cfa_stream_destroy(cfa_stream_t stream);
int cfa_device_sync(void);
/* GPU timers for the benchmark: events recorded on `stream` */
int cfa_timer_create(void** timer);
int cfa_timer_destroy(void* timer);
int cfa_timer_record(void* timer, cfa_stream_t stream);
int cfa_timer_elapsed_ms(void* t_begin, void* t_end, float* ms);
/* Work enqueued on `stream` after this call waits for `event` (a timer recorded on another
 * stream): the dependency between the SA-AMG set-up's two streams. */
int cfa_stream_wait(cfa_stream_t stream, void* event);

/* ---------------------------------------------------------------------------------------------
 * Mesh description on the device. Besides the reference's arrays (elem_v, x, y, bc_tag, CSR
 * graph; src/cfa/device.cuh:21-68) it carries the node -> incident-element table used by the
 * atomic-free owner-computes assembly: slot s in [inc_ptr[i], inc_ptr[i+1]) of node i holds
 * inc_elem[s] = 3 e + a (node i is vertex a of element e) and inc_pos[3 s + b] = CSR position
 * of entry (i, v_b(e)) in row i.
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int n_nodes, n_elems, nnz;
    const int* elem_v;   /* [3 n_elems] */
    const double* x;     /* [n_nodes] */
    const double* y;
    const int* bc_tag;   /* [n_nodes], BoundaryTag values of src/cfa/common.cuh:68-74 */
    const int* row_ptr;  /* [n_nodes + 1] */
    const int* col_ind;  /* [nnz] */
    const int* inc_ptr;  /* [n_nodes + 1] */
    const int* inc_elem; /* [3 n_elems] */
    const int* inc_pos;  /* [9 n_elems] */
    /* Outlet condition (src/cfa/common.cuh:76-89, device.cuh:153-172). outlet_dirichlet = 0 (the
     * default, also the zero-initialised struct): p = 0 on outlet nodes, nu du/dn = 0. Otherwise
     * the outlet carries the inflow parabola times outlet_scale, and the pressure increment is
     * Dirichlet (dp = 0) at node p_pin only. */
    int outlet_dirichlet;
    int p_pin;
    double outlet_scale;
} cfa_mesh_t;

/* Momentum system (src/cfa/fem_kernels.cuh:40-181 k_assemble_momentum_fem, fused with
 * :411-453 k_apply_bc_momentum). Writes A_val, rhs_u, rhs_v and mass (no prior zeroing), then
 * Dirichlet rows (inlet/wall/cylinder, and the outlet with m->outlet_dirichlet) become identity
 * rows whose rhs and u, v entries hold the boundary value: the inflow parabola with peak u_max
 * (times m->outlet_scale on the outlet), zero on walls and cylinder, or the manufactured solution
 * at mms_time when mms_time >= 0. u_n, v_n must not alias u, v.
 * fx, fy (nodal body force) may be NULL. */
int cfa_assemble_momentum(const cfa_mesh_t* m, const double* u_n, const double* v_n,
                          const double* u_prev, const double* v_prev, const double* p,
                          double nu, double rho, double a0, double a1, double a2, double c1,
                          double c2, double tau_time, double skew, const double* fx,
                          const double* fy, double u_max, double mms_time, double* A_val,
                          double* rhs_u, double* rhs_v, double* mass, double* u, double* v,
                          cfa_stream_t stream);

/* Manufactured-solution body force at nodes (src/cfa/fem_kernels.cuh:348-355). */
int cfa_mms_forcing(int n, const double* x, const double* y, double t, double* fx, double* fy,
                    cfa_stream_t stream);

/* OSS weights s_K = rho tau_K / dt (src/cfa/fem_kernels.cuh:697-719). */
int cfa_oss_weights(const cfa_mesh_t* m, const double* u, const double* v, double dt, double rho,
                    double nu, double stab, double* elem_s, cfa_stream_t stream);

/* G(f)_i = sum_K w_K |K|/3 grad f|_K, w_K = sqrt(elem_s[K]) or 1 if elem_s is NULL
 * (src/cfa/fem_kernels.cuh:726-757). Writes gx, gy (no prior zeroing). */
int cfa_scalar_gradient(const cfa_mesh_t* m, const double* f, const double* elem_s, double* gx,
                        double* gy, cfa_stream_t stream);

/* Lumped P1 mass m_i = sum_K |K|/3 (src/cfa/fem_kernels.cuh:890-909). Writes mass. */
int cfa_lumped_mass(const cfa_mesh_t* m, double* mass, cfa_stream_t stream);

/* Stabilised pressure Poisson system (src/cfa/fem_kernels.cuh:201-282
 * k_assemble_pressure_poisson, fused with the outlet Dirichlet rows and the symmetric column
 * elimination of :460-490 k_apply_bc_pressure, with the OSS operator of :807-852
 * k_assemble_oss_operator / k_inverse_lumped_mass when oss_val is not NULL). Writes Kp, rhs,
 * div_num, oss_val [Bx|By|BTx|BTy] (4 nnz) and inv_m; sets dp = 0 on Dirichlet rows. gpx, gpy
 * hold G(p^n) (cfa_scalar_gradient with elem_s). stab = 0 gives the plain PPE (elem_s, gpx,
 * gpy unused). */
int cfa_assemble_pressure(const cfa_mesh_t* m, const double* u_star, const double* v_star,
                          const double* p, const double* gpx, const double* gpy,
                          const double* mass, const double* elem_s, double dt, double rho,
                          double stab, double* Kp, double* rhs, double* div_num, double* dp,
                          double* oss_val, double* inv_m, cfa_stream_t stream);

/* With m->outlet_dirichlet, cfa_assemble_pressure leaves rhs[p_pin] at its assembled value and
 * this call completes the boundary condition (src/cfa/fem_kernels.cuh:493-531 k_sum,
 * k_subtract_mean and the pinned row of k_apply_bc_pressure): rhs -= (sum rhs / sum mass) mass,
 * then rhs[p_pin] = 0. Each sum adds blocks of 256 entries by the reference's in-block tree and
 * the block results in ascending block order (one of the orders its atomics allow). ws: at least
 * cfa_pressure_mean_ws_doubles(n) doubles. */
size_t cfa_pressure_mean_ws_doubles(int n);
int cfa_pressure_remove_mean(const cfa_mesh_t* m, const double* mass, double* rhs, double* ws,
                             cfa_stream_t stream);

/* cfa_pressure_remove_mean without a workspace argument and without any allocation. No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The block sums go to a
 * fixed device array of the library, 512 KiB per GPU, held until the process ends (its address is
 * looked up once per device, by cfa_init or by the first call), so meshes of at most
 * CFA_MEAN_NOALLOC_MAX_NODES nodes are accepted; a larger one returns CFA_E_SIZE before any device
 * work. Results are bit-identical to cfa_pressure_remove_mean's. Every call on a GPU uses the same
 * array: calls that overlap on one GPU (other streams or threads) race. Nothing in the array needs
 * clearing (each call writes every sum it reads). */
#define CFA_MEAN_NOALLOC_MAX_NODES 8388608
int cfa_pressure_remove_mean_noalloc(const cfa_mesh_t* m, const double* mass, double* rhs, cfa_stream_t stream);

/* OSS operator alone (src/cfa/device.cuh:304-311 assemble_oss_projection: fem_kernels.cuh:807-852
 * k_assemble_oss_operator and k_inverse_lumped_mass): writes oss_val [Bx|By|BTx|BTy] and inv_m
 * from the weights elem_s and the lumped mass. */
int cfa_assemble_oss_operator(const cfa_mesh_t* m, const double* elem_s, const double* mass,
                              double* oss_val, double* inv_m, cfa_stream_t stream);

/* out -= B^T M_L^{-1} B in on non-Dirichlet rows (src/cfa/fem_kernels.cuh:855-885,
 * src/cfa/device.cuh:316-323). gx, gy: n-vector scratch. */
int cfa_oss_apply(const cfa_mesh_t* m, const double* oss_val, const double* inv_m,
                  const double* in, double* out, double* gx, double* gy, cfa_stream_t stream);

/* The same operator in matrix-free element form (src/cfa/device.cuh:287-297 with
 * src/cfa/fem_kernels.cuh:763-797), for checking the assembled form. */
int cfa_oss_apply_matrix_free(const cfa_mesh_t* m, const double* elem_s, const double* mass,
                              const double* in, double* out, double* gx, double* gy,
                              cfa_stream_t stream);

/* out = c K f, K the P1 stiffness matrix (src/cfa/fem_kernels.cuh:627-649, applied to a zeroed
 * output as the driver does at src/cylinder_fem_amr.cu:467-472). */
int cfa_stiffness_apply(const cfa_mesh_t* m, const double* f, double c, double* out,
                        cfa_stream_t stream);

/* Velocity correction and pressure update of src/cfa/fem_kernels.cuh:536-582
 * (k_accumulate_grad_dp_and_vorticity) and :592-622 (k_velocity_pressure_correction):
 *   u -= (dt/rho) M^-1 G dp on interior rows and, unless m->outlet_dirichlet, outlet rows,
 *   p += p_update (dp - rot div_num / m),
 *   vort = M^-1 (N, omega_h(u*)). gx, gy, vort_num receive the nodal sums. */
int cfa_correction(const cfa_mesh_t* m, double dt, double rho, double p_update, double rot,
                   const double* div_num, const double* mass, const double* dp, double* u,
                   double* v, double* p, double* vort, double* gx, double* gy, double* vort_num,
                   cfa_stream_t stream);

/* Forces and probes of one step (src/cfa/fem_kernels.cuh:1046-1123 k_volume_forces,
 * :1129-1139 k_sample_pressure_probes, :657-692 k_surface_forces). diag[6] receives
 * [F_x, F_y volume brackets, p_front, p_back, F_x, F_y surface]; probe = {a0, b0, a1, b1}
 * (host array) with weights w0, w1; cyl_edges = [3 n_cyl] (a, b, owning element); layer =
 * the n_layer elements with a cylinder vertex (the only ones with a non-zero volume term;
 * NULL = scan all elements); ws: 8 * 1024 doubles. */
int cfa_forces(const cfa_mesh_t* m, const int* cyl_edges, int n_cyl, const int* layer,
               int n_layer, const double* u, const double* v, const double* u_old,
               const double* v_old, const double* u_prev, const double* v_prev, const double* p,
               double b0, double b1, double b2, double nu, double rho, const int* probe, double w0,
               double w1, double* diag, double* ws, cfa_stream_t stream);

/* AMR indicators (src/cfa/device.cuh:265-281 with src/cfa/fem_kernels.cuh:921-1035): lumped
 * mass, recovered disturbance gradient grad_rec [4 n], elem_error [2 ne] = [eta_v | eta_i]. */
int cfa_error_indicator(const cfa_mesh_t* m, const double* u, const double* v, double u_max,
                        double* mass, double* grad_rec, double* elem_error, cfa_stream_t stream);

/* Squared MMS error and norm integrals (src/cfa/fem_kernels.cuh:362-405): out[6], summed in a
 * fixed order. ws: 6 * 1024 doubles. */
int cfa_mms_errors(const cfa_mesh_t* m, const double* u, const double* v, const double* p,
                   double t, double* out, double* ws, cfa_stream_t stream);

/* ---------------------------------------------------------------------------------------------
 * Sparse linear algebra (src/cfa/linalg.cuh)
 * ------------------------------------------------------------------------------------------- */
/* y = A x (linalg.cuh:8-26); nrhs vectors of stride n (linalg.cuh:572-581 for nrhs = 2). */
int cfa_spmv(int n, const int* rp, const int* ci, const double* val, const double* x, double* y,
             int nrhs, cfa_stream_t stream);
/* Synchronous dot product, the reference's GpuLinearSolver::dot (linalg.cuh:172-177).
 * ws: 1024 doubles. */
int cfa_dot(int n, const double* x, const double* y, double* result, double* ws,
            cfa_stream_t stream);
/* BLAS-1 of linalg.cuh:50-90: y += a x; y = x + b y; out = a .* b. inv_diag = 1 / a_ii, or 1
 * where |a_ii| <= eps (eps 1e-14: linalg.cuh:28-48; eps 1e-300: the AMG rule of :871-878). */
int cfa_axpy(int n, double a, const double* x, double* y, cfa_stream_t stream);
int cfa_xpay(int n, const double* x, double b, double* y, cfa_stream_t stream);
int cfa_vec_mul(int n, const double* a, const double* b, double* out, cfa_stream_t stream);
int cfa_inv_diag(int n, const int* rp, const int* ci, const double* val, double eps,
                 double* inv_diag, cfa_stream_t stream);

/* OSS operator for the fused PCG apply. */
typedef struct {
    const double* oss_val; /* [4 nnz] */
    const double* inv_m;   /* [n] */
    const int* bc_tag;     /* [n] */
    double* gx;            /* [n] scratch */
    double* gy;
    int outlet_dirichlet;  /* as in cfa_mesh_t: skipped rows are node p_pin, else outlet nodes */
    int p_pin;
} cfa_oss_op_t;

/* One SA-AMG level (src/cfa/linalg.cuh:939-957 Level): operator A (rows x rows), prolongator
 * P (rows x cols), restriction R = P^T, inverse diagonal and Jacobi weight, work vectors.
 * tpr_*: lanes per row of the fused CSR kernels (1..64). */
typedef struct {
    int rows, cols;
    int a_nnz, p_nnz, r_nnz;
    int tpr_a, tpr_p, tpr_r;
    const int* a_rp; const int* a_ci; const double* a_v;
    const int* p_rp; const int* p_ci; const double* p_v;
    const int* r_rp; const int* r_ci; const double* r_v;
    const double* dinv;
    double omega;
    double* x; double* b; double* r; double* ax;
    /* FP32 copies of the values and FP32 work vectors for the mixed-precision cycle
     * (cfa_amg_t.use_f32); NULL when the hierarchy has none */
    const float* a_v32; const float* p_v32; const float* r_v32; const float* dinv32;
    float* x32; float* b32; float* r32; float* ax32;
    /* Sliced-ELL copy of A (cfa_sell_plan / cfa_sell_fill) that the level's smoothing sweeps
     * and residual read instead of the CSR arrays (the last sweep of a V-cycle, which emits the
     * r.z partials, stays on the CSR rows). Each row's products are summed in the order of the
     * CSR kernels with tpr_a lanes (1, 2 or 4), so the results are theirs bit for bit. s_sp is NULL
     * for none; s_len is its longest row (at most CFA_SELL_MAX_LEN); s_v32 holds its FP32 values
     * (use_f32). */
    int s_len;
    const int* s_sp; const int* s_row; const int* s_ci; const double* s_v; const float* s_v32;
} cfa_amg_level_t;

/* SA-AMG hierarchy for the V(ss, ss) cycle (src/cfa/linalg.cuh:1161-1206).
 * levels: host array of n_levels descriptors; coarse_inv: dense inverse (coarse_n x coarse_n,
 * row-major) of the coarsest operator; cb, cx: [coarse_n]. use_f32: run the cycle on the FP32
 * copies (smoothers, residuals, transfers in FP32; the coarse solve in FP64; r in and z out stay
 * FP64), with cb32, cx32 its coarse vectors. */
typedef struct {
    int n_levels;
    const cfa_amg_level_t* levels;
    int coarse_n;
    const double* coarse_inv;
    double* cb; double* cx;
    int smooth_steps;
    int use_f32;
    float* cb32; float* cx32;
} cfa_amg_t;

/* z = M^{-1} r, one V-cycle (src/cfa/linalg.cuh:1161-1169 AmgPreconditioner::apply). */
int cfa_amg_apply(const cfa_amg_t* amg, const double* r, double* z, cfa_stream_t stream);

/* out[i] = (float) in[i], i < n: the FP32 copies of a hierarchy. */
int cfa_to_f32(size_t n, const double* in, float* out, cfa_stream_t stream);

/* One damped Jacobi step on level `level`: y = x + omega D^{-1} (b - A x), the smoother of the
 * V-cycle (src/cfa/linalg.cuh:1177-1181 jacobi_step); one kernel. */
int cfa_amg_smooth(const cfa_amg_t* amg, int level, const double* x, const double* b, double* y,
                   cfa_stream_t stream);
/* The same step in FP32 on the level's FP32 copies (amg->use_f32; the mixed-precision cycle's
 * smoother); CFA_E_NOTREADY without them. */
int cfa_amg_smooth_f32(const cfa_amg_t* amg, int level, const float* x, const float* b, float* y,
                       cfa_stream_t stream);

/* Sliced-ELL copy of a CSR operator for the V-cycle's sweeps (no counterpart in the reference,
 * whose V-cycle reads the CSR arrays, linalg.cuh:884-901; the sums are the CSR kernels'). Positions come in slices of 64 (one wavefront). Inside each window of 256
 * positions the rows are ordered by decreasing length, ties by row index, so the rows of a slice
 * have nearly the same length. The slice length len_q is the length of its first row. Entry k of
 * the row at position 64 q + l lies at sell_sp[q] + 64 k + l; slots past a row's end hold column 0
 * and value 0.
 * cfa_sell_plan writes:
 *   sell_row [64 * slices], slices = ceil(rows / 64): the row at each position, -1 past the last row;
 *   sell_sp [slices + 1]: the slice offsets;
 *   info (device int [2]): the entries of the copy and the longest row.
 * Its ws holds cfa_sell_ws_bytes(rows) bytes.
 * cfa_sell_fill then writes sell_ci and sell_v [info[0]].
 * The V-cycle takes the copy when the longest row is at most CFA_SELL_MAX_LEN and tpr_a is 1, 2
 * or 4. */
#define CFA_SELL_MAX_LEN 16
size_t cfa_sell_ws_bytes(int rows);
int cfa_sell_plan(int rows, const int* rp, int* sell_row, int* sell_sp, int* info, void* ws, size_t ws_bytes,
                  cfa_stream_t stream);
int cfa_sell_fill(int rows, const int* rp, const int* ci, const double* val, const int* sell_row,
                  const int* sell_sp, int* sell_ci, double* sell_v, cfa_stream_t stream);

/* Host callback that enqueues out = op(in) (or out += op(in) for an extra operator) on the
 * stream; the generic form of the std::function arguments of GpuLinearSolver::solve_pcg. */
typedef void (*cfa_op_cb)(const double* in, double* out, void* ctx);

/* Krylov workspace: device vectors of capacity >= n (x 2 for the two-component solver),
 * device scalars, and a pinned status word the kernels update so the host can follow the
 * iteration without a synchronisation per dot product. Layout sizes: vec[12] vectors of
 * 2 * cap doubles, part: 8 * 4096 doubles, scal: 64 doubles, status: 16 ints (host, mapped).
 * internal: NULL initially; the library keeps its captured iteration graphs there (released by
 * cfa_krylov_ws_release). use_graphs: replay each iteration as a captured HIP graph (set to 0
 * to launch the kernels one by one; solves with host callbacks never use graphs). */
typedef struct {
    int cap;
    double* vec[12];
    double* part;
    double* scal;
    int* status_host;
    int* status_dev;
    void* internal;
    int use_graphs;
} cfa_krylov_ws_t;

/* Frees what the library attached to ws->internal (captured graphs, its capture stream). */
int cfa_krylov_ws_release(cfa_krylov_ws_t* ws);

typedef struct {
    int iters[2];
    int converged[2];
    double rel_residual[2];
} cfa_solve_result_t;

/* Preconditioned CG of GpuLinearSolver::solve_pcg (src/cfa/linalg.cuh:183-256): warm start
 * from x, stop when ||r|| < tol ||b||, x = 0 if ||b|| < 1e-14, breakdown if |p.Ap| < 1e-30.
 * Operator: CSR val (+ oss or extra_cb); preconditioner: amg, else prec_cb, else Jacobi.
 * Scalars stay on the device; the host only follows the status word. */
int cfa_pcg(int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
            int max_iter, double tol, const cfa_oss_op_t* oss, cfa_op_cb extra_cb,
            void* extra_ctx, const cfa_amg_t* amg, cfa_op_cb prec_cb, void* prec_ctx,
            cfa_krylov_ws_t* ws, cfa_solve_result_t* result, cfa_stream_t stream);

/* out = (Kp + OSS) in, the operator of the fused PCG iteration (CSR SpMV of
 * src/cfa/linalg.cuh:199-203 plus apply_oss_projection of src/cfa/device.cuh:310-323) in its
 * two kernels: stage 1 computes out = Kp in and g = M^-1 B in (oss->gx, oss->gy), stage 2
 * out -= B^T g on non-outlet rows and writes in.out block partials to part [4096];
 * stage 0 runs both. */
int cfa_pcg_operator(int n, const int* rp, const int* ci, const double* val, const cfa_oss_op_t* oss,
                     const double* in, double* out, double* part, int stage, cfa_stream_t stream);

/* Multicolour ILU(0) (src/cfa/linalg.cuh:361-568). colour_start is a host array. */
typedef struct {
    int n, nnz, n_colours, nl, nu;
    const int* colour_start; /* host [n_colours + 1] */
    const int* rows;         /* [n] rows in colour order */
    const int* colour;       /* [n] */
    const int* diag_pos;     /* [n] CSR position of the diagonal of row i */
    const int* row_ptr; const int* col_ind;
    double* lu;              /* [nnz] factor in CSR layout */
    const int* lp; const int* up;       /* [n + 1] split pattern, colour-row order: L holds the
                                           columns of earlier colours (none in colour 0), U those
                                           of later colours (none in the last) */
    const int* lc; const int* uc;       /* [nl], [nu] column of each split entry, as its colour-order
                                           position (pos[column]) */
    const int* pos;                     /* [n] colour-order position of row i (rows[pos[i]] = i) */
    double* r_perm; double* z_perm;     /* [2 n] work vectors of the apply in colour order */
    const int* l_src; const int* u_src; /* CSR position of each split entry */
    const int* d_src;                   /* [n] CSR position of the diagonal, colour-row order */
    double* lv; double* uv; double* dinv;
    float* lv32; float* uv32; float* dinv32; /* FP32 copies (mixed precision); may be NULL */
    /* elimination schedule of the factorization, per colour-row t (colour by colour, CSR order
     * inside a colour, as linalg.cuh:366-392): el_ptr [n + 1]; el [4 per step] = (CSR position
     * of l_ik, CSR position of u_kk, first and end update); upd [2 per update] = (position of
     * a_ij, position of u_kj): a_ij -= l_ik u_kj */
    const int* el_ptr; const int* el; const int* upd;
    /* Sliced-ELL copies of the split rows for the one-lane colour sweeps: the FP64 forward and
     * backward sweeps and the FP32 backward sweep (cfa_sell_plan's layout, built per colour).
     * NULL s_lrow / s_urow selects the CSR rows. Colour c's rows take positions s_pos0[c] + t,
     * t < 64 ceil(m_c / 64); s_pos0 is a host array [n_colours + 1], each entry a multiple of 256.
     * Inside 256-position windows the rows are ordered by decreasing length, ties by row.
     *   s_lrow / s_urow: the colour's row t at each position, -1 past its last;
     *   s_lsp / s_usp: slice offsets; entry k of the row at position p lies at sp[p / 64] + 64 k + p % 64;
     *   s_lc / s_uc: colour-order columns (0 in padding);
     *   s_lsrc / s_usrc: the index into lv / uv of each slot (-1 in padding);
     *   s_lv / s_uv and their FP32 copies: the values, refreshed by cfa_ilu0_factor;
     *   s_llen / s_ulen: the longest L / U row (at most CFA_SELL_MAX_LEN).
     * The sweeps sum each row in its CSR order, and padding adds exact zeros, so the results are
     * the CSR kernels'. */
    const int* s_pos0;
    const int* s_lrow; const int* s_lsp; const int* s_lc; const int* s_lsrc; double* s_lv; float* s_lv32;
    const int* s_urow; const int* s_usp; const int* s_uc; const int* s_usrc; double* s_uv; float* s_uv32;
    int s_llen, s_ulen, s_lslots, s_uslots;
} cfa_ilu_t;

/* Factor the values (linalg.cuh:531-541): lu = ILU(0) of val in colour order, then lv, uv,
 * dinv (and the FP32 copies when present). */
int cfa_ilu0_factor(const cfa_ilu_t* ilu, const double* val, cfa_stream_t stream);
/* z = (LU)^{-1} r for nrhs vectors of stride n (linalg.cuh:544-555). The sweeps run in colour
 * order: r is permuted into r_perm, every colour's rows, factor entries and vector entries are
 * then contiguous, and z is permuted back; each row's arithmetic is the reference's, in its
 * order. */
int cfa_ilu0_apply(const cfa_ilu_t* ilu, const double* r, double* z, int nrhs,
                   cfa_stream_t stream);

/* Two-component BiCGSTAB with the multicolour ILU(0) (GpuBicgstab2::solve,
 * src/cfa/linalg.cuh:660-754): per-component coefficients and stopping rule
 * ||r|| < tol ||r0|| or < 1e-11, breakdown guard 1e-30 ||r0||^2 + 1e-300; x_u, x_v are the
 * initial guesses and receive the solutions. use_f32: apply the FP32 copy of the factors. */
int cfa_bicgstab2(int n, const int* rp, const int* ci, const double* val, const double* b_u,
                  const double* b_v, double* x_u, double* x_v, int max_iter, double tol,
                  const cfa_ilu_t* ilu, int use_f32, cfa_krylov_ws_t* ws,
                  cfa_solve_result_t* result, cfa_stream_t stream);

/* Jacobi-preconditioned BiCGSTAB for one right-hand side (GpuLinearSolver::solve_bicgstab,
 * src/cfa/linalg.cuh:261-347). */
int cfa_bicgstab_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b,
                        double* x, int max_iter, double tol, cfa_krylov_ws_t* ws,
                        cfa_solve_result_t* result, cfa_stream_t stream);

/* ---- SA-AMG set-up on the device (src/cfa/linalg.cuh:995-1060, 1276-1431, host code in the
 * reference) -----------------------------------------------------------------------------------
 * Bit-identical to the reference's host set-up: the same operations in the same order
 * (contraction off), the greedy aggregation included (cfa_amg_aggregate).
 * cfa_amg_dinv: dinv = 1 / a_ii, or 1 where |a_ii| <= 1e-300 (linalg.cuh:1003-1008).
 * cfa_amg_power: spectral_radius_DinvA (linalg.cuh:1276-1300): `iterations` power steps of
 * D^-1 A from the device vector x0 (the reference starts from 1 + 0.1 sin(1 + i)), norms summed
 * per chunk of 4096 rows in row order and over the chunks in order; *lambda (host) = the last
 * ||D^-1 A x|| / ||x||. ws: cfa_amg_power_ws_doubles(n) doubles.
 * cfa_spgemm: C = A B in two passes (csr_multiply, linalg.cuh:799-844): pass 0 writes the
 * number of distinct columns of each row to cnt and sets *overflow (device int, cleared by the
 * caller) when a row is beyond the method's limit; pass 1 writes the sorted columns and the
 * values, each summed from 0 in the order the products are met, at c_rp (from
 * cfa_scan_counts). method 0: one thread per row merging the right rows (rows of A with at
 * most 16 entries); method 1: one workgroup per row sorting up to 4096 products in LDS;
 * method 2: one wavefront per row merging the right rows staged in LDS (rows of A with at most
 * 64 entries and 1024 products). Rows of B must have sorted, unique columns.
 * cfa_scan_counts: out[0..n] = exclusive prefix sums of cnt (out[n] = total; also copied to the
 * device int *total when given); ws: device workspace of cfa_scan_counts_ws_bytes(n) bytes.
 * cfa_amg_power with lambda == NULL does not synchronize: {||D^-1 A x||, lambda} of the last step
 * are then left in ws[cfa_amg_power_ws_doubles(n) - 2 .. - 1].
 * cfa_amg_prolongator: P = P0 - omega D^-1 A P0 from AP0 = A P0 (smoothed_prolongator,
 * linalg.cuh:1360-1392: the row's own aggregate merged in column order, exact zeros dropped);
 * pass 0 counts per row, pass 1 fills at p_rp.
 * cfa_csr_transpose: T = M^T for a rows x cols CSR matrix with nnz entries (csr_transpose,
 * linalg.cuh:847-862: the entries of a column in ascending row order, values moved unchanged);
 * t_rp [cols + 1], t_ci / t_v [nnz]; ws: device workspace of cfa_csr_transpose_ws_bytes bytes. */
int cfa_amg_dinv(int n, const int* rp, const int* ci, const double* val, double* dinv, cfa_stream_t stream);
size_t cfa_amg_power_ws_doubles(int n);
int cfa_amg_power(int n, const int* rp, const int* ci, const double* val, const double* dinv, const double* x0,
                  int iterations, double* ws, double* lambda, cfa_stream_t stream);
int cfa_spgemm(int rows, const int* a_rp, const int* a_ci, const double* a_v, const int* b_rp, const int* b_ci,
               const double* b_v, int pass, int method, int* cnt, const int* c_rp, int* c_ci, double* c_v,
               int* overflow, cfa_stream_t stream);
/* cfa_spgemm_sorted: C = A B (csr_multiply, linalg.cuh:799-844) for rows beyond cfa_spgemm's
 * limits: every product is listed in encounter order under the key (row, column), stably radix
 * sorted, and each key's products are summed from 0 in that order (the same values as
 * cfa_spgemm). A is rows x (rows of B) with a_nnz entries, B has `cols` columns. Three passes on
 * one workspace: pass 0 (ws of cfa_spgemm_sorted_ws_bytes(rows, cols, a_nnz, 0) bytes) sets the
 * host *count = products; pass 1 (ws of cfa_spgemm_sorted_ws_bytes(..., products) bytes) lists and
 * sorts them and sets the host *count = entries of C; pass 2 on the same workspace writes c_rp
 * [rows + 1], c_ci and c_v [entries]. Passes 0 and 1 synchronize the stream. At most 2^31 - 1
 * products (the size query returns 0 beyond). */
size_t cfa_spgemm_sorted_ws_bytes(int rows, int cols, int a_nnz, long long products);
int cfa_spgemm_sorted(int rows, int cols, int a_nnz, const int* a_rp, const int* a_ci, const double* a_v,
                      const int* b_rp, const int* b_ci, const double* b_v, int pass, long long products, void* ws,
                      size_t ws_bytes, long long* count, int* c_rp, int* c_ci, double* c_v, cfa_stream_t stream);
size_t cfa_scan_counts_ws_bytes(int n);
int cfa_scan_counts(int n, const int* cnt, int* out, int* total, void* ws, size_t ws_bytes, cfa_stream_t stream);
int cfa_amg_prolongator(int n, const int* ap_rp, const int* ap_ci, const double* ap_v, const int* agg,
                        const double* p0v, const double* dinv, double omega, int pass, int* cnt, const int* p_rp,
                        int* p_ci, double* p_v, cfa_stream_t stream);
size_t cfa_csr_transpose_ws_bytes(int rows, int cols, int nnz);
/* cfa_amg_aggregate: the three-phase greedy aggregation of smoothed_prolongator (linalg.cuh:
 * 1306-1356) for the n x n CSR matrix (rp, ci, val) with inverse diagonal dinv and strength
 * threshold theta: agg [n] (aggregate of each row, numbered as the row-order loop numbers them),
 * p0v [n] = 1 / sqrt(size of the row's aggregate), info (device int [4]) = aggregates, phase-1
 * aggregates, longest row, rounds of the decision front. ws: cfa_amg_aggregate_ws_bytes bytes. */
size_t cfa_amg_aggregate_ws_bytes(int n, int nnz);
int cfa_amg_aggregate(int n, int nnz, const int* rp, const int* ci, const double* val, const double* dinv,
                      double theta, void* ws, size_t ws_bytes, int* agg, double* p0v, int* info, cfa_stream_t stream);
int cfa_csr_transpose(int rows, int cols, int nnz, const int* rp, const int* ci, const double* v, void* ws,
                      size_t ws_bytes, int* t_rp, int* t_ci, double* t_v, cfa_stream_t stream);

/* ---- mesh topology on the device (src/cfa/mesh.cuh:198-376, host code in the reference) -------
 * cfa_mesh_topology: from the element list elem_v [3 ne] of a mesh with n nodes, the node-element
 * incidence (inc_ptr [n + 1]; inc_elem [3 ne] = 3 e + corner, each node's elements ascending),
 * the CSR graph (row i = {i} and the other vertices of its elements, sorted: row_ptr [n + 1],
 * col_ind with room for n + 6 ne entries) and the owner-computes position table inc_pos [9 ne]
 * (CSR position of each element vertex in the row of each incidence), as
 * build_csr_and_boundary_topology (mesh.cuh:198-270) and the incidence build of the kernels;
 * *nnz_out (device int) = the graph's entries, or -1 when a node has more than 47 elements (the
 * caller then builds on the host).
 * cfa_mesh_check: validate_mesh (mesh.cuh:283-376): stats (device) [6] = non-positive elements,
 * element edges missing from the graph, non-conforming edges (shared by more than two
 * elements, or by one off a boundary piece of the channel h_channel x l_channel or the
 * cylinder), minimum signed area, total area (summed in a fixed tree order), maximum corner
 * cosine (the minimum angle is its arccosine).
 * ws: device workspace of cfa_topology_ws_bytes(n, ne) bytes, shared by both calls. */
size_t cfa_topology_ws_bytes(int n_nodes, int n_elems);
/* cfa_sort_pairs_u32: stable ascending sort of n (key, value) pairs by the 32-bit key (ties keep
 * their input order); the Morton-curve ordering of the leaves in AdaptiveMesh::extract
 * (src/cfa/amr.cuh:152-228, a std::sort of (key, id) on the host there) when the ids come in
 * ascending order. ws: cfa_sort_pairs_ws_bytes(n) bytes. */
size_t cfa_sort_pairs_ws_bytes(int n);
int cfa_sort_pairs_u32(int n, const unsigned* keys_in, const int* vals_in, unsigned* keys_out, int* vals_out,
                       void* ws, size_t ws_bytes, cfa_stream_t stream);
int cfa_mesh_topology(int n, int ne, const int* elem_v, void* ws, size_t ws_bytes, int* inc_ptr, int* inc_elem,
                      int* row_ptr, int* col_ind, int* inc_pos, int* nnz_out, cfa_stream_t stream);
/* cfa_mesh_incidence: the incidence and position tables of cfa_mesh_topology for a graph the
 * caller already has (row_ptr, col_ind on the device). */
int cfa_mesh_tables(int n, int ne, const int* elem_v, const int* row_ptr, const int* col_ind, void* ws,
                    size_t ws_bytes, int* inc_ptr, int* inc_elem, int* inc_pos, cfa_stream_t stream);
int cfa_mesh_check(int n, int ne, const int* elem_v, const double* x, const double* y, const int* bc_tag,
                   const int* row_ptr, const int* col_ind, int nnz, double h_channel, double l_channel, void* ws,
                   size_t ws_bytes, double* stats, cfa_stream_t stream);

/* ---- newest-vertex-bisection forest on the device (src/cfa/amr.cuh:58-996, host code in the
 * reference: AdaptiveMesh's forest, marking, look-ahead, coarsening, refinement with conformity
 * closure, Morton extraction and solution transfer) -----------------------------------------------
 * cfa_forest_t: the forest in structure-of-arrays form. Element slot e: ev[3e..3e+2] = newest
 * vertex, refinement edge (v1, v2), counter-clockwise; egen, eparent, ech[2e..2e+1] (-1: leaf),
 * ealive. Vertex slot v: coordinates, the five solution values (u, v, p and the previous time
 * level uo, vo), boundary tag, valive, vpar[2v..2v+1] (the elements bisected through v), vedge (the
 * bisected edge, min id in the high word; 0 for a base vertex). efree / vfree: recycled slots,
 * taken from the end (LIFO, amr.cuh:761-795). hkey / hval: open-addressing table bisected edge ->
 * midpoint (hash_cap a power of two, load kept below one half by the caller). log [3 x]: the
 * midpoints created by the current adaptation {m, a, b} in creation order; cyl_log: those on the
 * cylinder, whose arc positions the caller may set itself (cfa_nvb_set_xy) before the next pass
 * uses them. cnt [CFA_NVB_COUNTERS]
 * (device): the counters below, read and written by the kernels so a whole adaptation runs without
 * host round trips except one per refinement pass. Capacities are the caller's: before a call the
 * arrays must hold the slots it may create (bisect: 2 n elements, n vertices, n log entries). */
enum {
    CFA_NVB_NE = 0,        /* element slots in use (alive or recycled) */
    CFA_NVB_NV,            /* vertex slots in use */
    CFA_NVB_EFREE,         /* entries of efree */
    CFA_NVB_VFREE,         /* entries of vfree */
    CFA_NVB_LOG,           /* entries of log */
    CFA_NVB_ERROR,         /* non-zero: an edge bisected by a third element */
    CFA_NVB_LIST,          /* length of the last selected list */
    CFA_NVB_MARKED_REFINE,
    CFA_NVB_MARKED_COARSEN,
    CFA_NVB_REMOVED,       /* vertices removed by coarsening */
    CFA_NVB_NODES,         /* nodes of the last extraction */
    CFA_NVB_ORDERED,       /* leaves above tolerance in the marking order */
    CFA_NVB_CANDIDATES,    /* vertices collapsed by the last coarsening sweep */
    CFA_NVB_CYLINDER,      /* entries of cyl_log */
    CFA_NVB_COUNTERS = 16
};
typedef struct {
    int* ev;
    int* egen;
    int* eparent;
    int* ech;
    unsigned char* ealive;
    double* vx;
    double* vy;
    double* vu;
    double* vv;
    double* vp;
    double* vuo;
    double* vvo;
    int* vtag;
    unsigned char* valive;
    int* vpar;
    unsigned long long* vedge;
    int* efree;
    int* vfree;
    int* log;
    int* cyl_log;   /* [3 x]: {m, a, b} of the cylinder midpoints created (any order) */
    unsigned long long* hkey;
    int* hval;
    int hash_cap;
    int* cnt;
} cfa_forest_t;

/* Marking phase of one adaptation (amr.cuh:238-455 up to the coarsening sweeps): indicators
 * normalised by their tolerances (optionally grown with the distance from the cylinder), leaf edge
 * adjacency, advective look-ahead (traces along the centroid velocity with a walk across edges and
 * a 0.01 bucket-grid fallback; a leaf keeps the maximum of what reaches it), the refinement marks
 * largest score first within the element budget (ties in leaf order), the minimum level, two
 * hole-filling passes, buffer layers, the protected layer and the coarsening flags. Leaves are the
 * last extraction's, in its order. Forest-indexed outputs av, ai, target, refine, coarsen cover the
 * element slots; clist receives the coarsening-flagged leaves in leaf order (count in
 * cnt[CFA_NVB_LIST]). */
typedef struct {
    int n_leaves;                  /* L */
    const int* leaves;             /* [L] forest slot of each leaf */
    const double* eta;             /* [2 L] vorticity, then interpolation indicator */
    int max_level, min_level, buffer_layers;
    long long max_elems;
    double tol_vort, tol_interp, coarsen_safety, lookahead_time, keep_radius, tol_radius, tol_power;
    double* av;                    /* [element slots] */
    double* ai;
    int* target;
    unsigned char* refine;
    unsigned char* coarsen;
    unsigned char* vflag;          /* [vertex slots] */
    int* nbr;                      /* [3 L] */
    int* clist;                    /* [L] */
    long long* predicted;          /* device: the predicted element count of the marks */
} cfa_nvb_mark_t;

/* cfa_nvb_tmp_bytes: rocPRIM temporary storage for lists of up to n items (select, scan, sort).
 * cfa_nvb_scratch_ints: int scratch (device) the calls below need for up to n_elem element slots,
 * n_vert vertex slots and L leaves.
 * cfa_nvb_hash_build: empties the table and inserts the live midpoints (vertex slots < nv).
 * cfa_nvb_leaves: list (ascending slot) of the element slots < ne that are leaves (mode 0), leaves
 * below target[e] (mode 1: the work of a refinement round, amr.cuh:476-508) or hanging leaves
 * (mode 2: an edge with a midpoint, amr.cuh:922-969); count in cnt[CFA_NVB_LIST].
 * cfa_nvb_bisect: bisects the n listed leaves as if one after the other in list order
 * (amr.cuh:802-880): (c, a, b) -> (m, c, a), (m, b, c); the midpoint of (a, b) is taken from the
 * table or created (slot from vfree, else appended; cylinder midpoints on the circle, boundary
 * tags by position, P1 values with exact Dirichlet data, inlet peak inlet_umax; logged); children
 * take slots from efree, else appended; with inherit they take target[e], else -1.
 * cfa_nvb_mark: the marking phase above (scratch: cfa_nvb_scratch_ints ints).
 * cfa_nvb_coarsen_sweep: one sweep of vertex collapses (amr.cuh:391-441, 977-996): the newest
 * vertices of the listed leaves whose two (or one, on the boundary) bisected elements have
 * coarsening-flagged leaf children predicted below coarsen_safety are removed in ascending slot
 * order, the merged parents inherit the flag and the merged indicators (sqrt(2) and 2 times the
 * larger), freed slots go to the free lists in that order; clist is replaced by the merged
 * parents (count in cnt[CFA_NVB_LIST], removals added to cnt[CFA_NVB_REMOVED]).
 * cfa_nvb_extract: compact mesh of the listed leaves (ascending slot, count L): leaves sorted by the
 * Morton key of the centroid (16 bits per axis, ties by slot), nodes numbered by first use
 * (amr.cuh:152-228); out_leaf [L], out_node [nodes], elem_v [3 L], gen [L] and the node arrays;
 * node count in cnt[CFA_NVB_NODES].
 * cfa_nvb_uniform_targets: target = -1 on the ne element slots, the listed leaves' generation plus
 * 2 levels (refine_uniform, amr.cuh:461-471).
 * cfa_nvb_set_xy: vertex slots slots [n] take the coordinates xy [2 n] (x, y pairs).
 * cfa_nvb_scatter: values [5 n] (u, v, p, uo, vo of n nodes) to the vertex slots nodes [n].
 * cfa_nvb_replay: recomputes the midpoint values of log entries [first, last) (amr.cuh:112-145,
 * 896-914); entries of one refinement pass are independent, passes must be replayed in order.
 * cfa_nvb_gather: values [5 n] from the vertex slots nodes [n]. */
size_t cfa_nvb_tmp_bytes(int n);
size_t cfa_nvb_scratch_ints(int n_elem, int n_vert, int n_leaves);
int cfa_nvb_hash_build(const cfa_forest_t* f, int nv, cfa_stream_t stream);
int cfa_nvb_leaves(const cfa_forest_t* f, int ne, int mode, const int* target, int* flags, int* list, void* tmp,
                   size_t tmp_bytes, cfa_stream_t stream);
int cfa_nvb_bisect(const cfa_forest_t* f, int n, const int* list, int inherit, int* target, double inlet_umax,
                   int* scratch, void* tmp, size_t tmp_bytes, cfa_stream_t stream);
int cfa_nvb_mark(const cfa_forest_t* f, const cfa_nvb_mark_t* m, int ne, int nv, int* scratch, void* tmp,
                 size_t tmp_bytes, cfa_stream_t stream);
int cfa_nvb_coarsen_sweep(const cfa_forest_t* f, const cfa_nvb_mark_t* m, int nv, int* scratch, void* tmp,
                          size_t tmp_bytes, cfa_stream_t stream);
int cfa_nvb_extract(const cfa_forest_t* f, int nv, int n_leaves, const int* leaves, int* scratch, void* tmp,
                    size_t tmp_bytes, int* out_leaf, int* out_node, int* elem_v, int* gen, double* x, double* y,
                    int* tag, double* values, cfa_stream_t stream);
int cfa_nvb_uniform_targets(const cfa_forest_t* f, int ne, int n_leaves, const int* leaves, int levels, int* target,
                            cfa_stream_t stream);
int cfa_nvb_set_xy(const cfa_forest_t* f, int n, const int* slots, const double* xy, cfa_stream_t stream);
int cfa_nvb_scatter(const cfa_forest_t* f, int n, const int* nodes, const double* values, cfa_stream_t stream);
int cfa_nvb_replay(const cfa_forest_t* f, int first, int last, double inlet_umax, cfa_stream_t stream);
int cfa_nvb_gather(const cfa_forest_t* f, int n, const int* nodes, double* values, cfa_stream_t stream);

/* ---- discrete adjoint of the step (no counterpart in the reference) --------------------------
 * Vector-Jacobian products of the step's assemblies and updates (the forward operations are
 * those of src/cylinder_fem_amr.cu:277-538; the reference has no derivative code). A linear
 * solve x = A(q)^-1 b(q) is differentiated as q_bar = d/dq [lambda^T (b(q) - A(q) x)] with
 * A^T lambda = x_bar solved by the caller. All results are ADDED to the *_bar outputs and to
 * theta_bar[0] (viscosity) and theta_bar[1] (inflow peak velocity). x_bar, y_bar (node
 * coordinates, shape sensitivities) may both be NULL, which skips that part. ws: device
 * workspace of cfa_adj_workspace_doubles(n, ne) doubles. Deterministic: element contributions
 * are summed per node in element order. */
size_t cfa_adj_workspace_doubles(int n, int ne);
/* Momentum system (fem_kernels.cuh:40-181, :411-453) at x = (u_star, v_star) with adjoint
 * solutions lam_u, lam_v of A^T: sensitivities to u^n, v^n (u_n), u^{n-1}, v^{n-1} (u_prev),
 * p^n, the coordinates, viscosity and the inflow peak (inlet data 4 u_max y (H - y) / H^2). */
int cfa_adj_momentum(const cfa_mesh_t* m, const double* u_n, const double* v_n, const double* u_prev,
                     const double* v_prev, const double* p, const double* u_star, const double* v_star,
                     const double* lam_u, const double* lam_v, double nu, double rho, double a0, double a1,
                     double a2, double c1, double c2, double tau_time, double skew, double h_channel,
                     double* ws, double* u_n_bar, double* v_n_bar, double* u_prev_bar, double* v_prev_bar,
                     double* p_bar, double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream);
/* Pressure increment system (OSS weights fem_kernels.cuh:697-719, PPE :201-282 with the outlet
 * rows of :460-531, OSS operator :763-885) at the solution dp with adjoint solution lam of the
 * same (self-adjoint on the free rows) operator: sensitivities to u*, v*, p^n, the coordinates
 * (including those of the lumped mass in M^-1) and viscosity. */
int cfa_adj_pressure(const cfa_mesh_t* m, const double* elem_s, const double* u_star, const double* v_star,
                     const double* p, const double* dp, const double* lam, const double* mass, double nu, double rho,
                     double dt, double stab, double* ws, double* u_star_bar, double* v_star_bar, double* p_bar,
                     double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream);
/* Rotational pressure update and velocity correction (fem_kernels.cuh:536-649):
 * dp_bar = p1_bar + nu dt K M^-1 p1_bar - (dt/rho) G^T M^-1 u1_bar (corrected rows), zero on
 * outlet rows (overwritten, not added); theta_bar[0] += p1_bar . (-rho div_num / m); with
 * x_bar, y_bar the coordinate sensitivities of G, K and M at the increment dp. The sensitivities
 * u*_bar = u1_bar and p^n_bar = p1_bar are identities left to the caller. */
int cfa_adj_correction(const cfa_mesh_t* m, double dt, double rho, double nu, const double* div_num,
                       const double* mass, const double* dp, const double* p1_bar, const double* u1_bar,
                       const double* v1_bar, double* ws, double* dp_bar, double* x_bar, double* y_bar,
                       double* theta_bar, cfa_stream_t stream);
/* Volume-integral force (fem_kernels.cuh:1046-1123): gradient of wx F_x + wy F_y with respect
 * to u, v, u_old, v_old, u_prev, v_prev, p, the coordinates and viscosity. */
int cfa_adj_forces(const cfa_mesh_t* m, const double* u, const double* v, const double* u_old,
                   const double* v_old, const double* u_prev, const double* v_prev, const double* p, double b0,
                   double b1, double b2, double nu, double rho, double wx, double wy, double* ws, double* u_bar,
                   double* v_bar, double* u_old_bar, double* v_old_bar, double* u_prev_bar, double* v_prev_bar,
                   double* p_bar, double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream);
/* out[k] = in[perm[k]], k < nnz (CSR values of A^T on a symmetric pattern); x *= a. */
int cfa_gather_values(int nnz, const int* perm, const double* in, double* out, cfa_stream_t stream);
int cfa_scale(int n, double a, double* x, cfa_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CFA_HIP_H */
