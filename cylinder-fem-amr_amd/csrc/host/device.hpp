// Device state of one mesh and the GPU driver routines shared by time stepping and remeshing.
// Interfaces of src/cfa/device.cuh:21-574 (DeviceMeshBuffers, compute_error_indicator,
// apply_oss_projection[_matrix_free], assemble_oss_projection, project_scheme_consistent,
// project_divergence_free, set_initial_condition, print_amr_report, RemeshProfile) and of the
// cylinder probes of src/cfa/diagnostics.cuh:8-45.
#pragma once

#include <functional>
#include <string>

#include "amr.hpp"
#include "gpu.hpp"
#include "linalg.hpp"

namespace cfa {

struct PressureProbe {
    int a = 0, b = 0;
    double w = 0.0;   // p = (1 - w) p_a + w p_b
};
PressureProbe find_cylinder_probe(const Mesh2D& mesh, double theta);

struct DeviceMeshBuffers {
    int n_nodes = 0, n_elems = 0, nnz = 0;
    int n_cyl_edges = 0, n_layer = 0;

    DevArray<int> d_elem_v, d_cyl_edges, d_bc_tag, d_row_ptr, d_col_ind;
    DevArray<int> d_inc_ptr, d_inc_elem, d_inc_pos, d_layer;   // owner-computes tables
    DevArray<double> d_x, d_y, d_u, d_v, d_u_old, d_v_old, d_u_prev, d_v_prev, d_p, d_dp, d_vort;
    DevArray<double> d_A_val, d_Kp_val, d_rhs_u, d_rhs_v, d_rhs_p, d_mass_lumped;
    DevArray<double> d_grad_dp_x, d_grad_dp_y, d_vort_num, d_elem_error, d_grad_rec, d_div_num;
    DevArray<double> d_elem_s, d_oss_gx, d_oss_gy, d_diag, d_oss_val, d_inv_m, d_phi, d_ws;

    PressureProbe probe_front, probe_back;

    // Outlet condition (src/cfa/common.cuh:76-89; the reference's global g_outlet_dirichlet, set
    // by the program from --outlet). With outlet_dirichlet every upload picks the pinned node and
    // the outlet parabola factor (device.cuh:153-172) and flux-balances the transferred outlet
    // velocities (:179-199).
    bool outlet_dirichlet = false;
    int p_pin = -1;
    double outlet_scale = 1.0;

    const cfa_mesh_t* view() const { return &view_; }
    cfa_oss_op_t oss_op() const;
    // Mean removal and pinned row of the pure-Neumann PPE after cfa_assemble_pressure
    // (src/cfa/fem_kernels.cuh:519-531 apply_bc_pressure); nothing with the default outlet.
    void pressure_outlet_bc(const double* mass, double* rhs, cfa_stream_t s);

    // Upload of a mesh whose CSR graph the host built (build_csr_and_boundary_topology).
    void sync_from_host(const Mesh2D& mesh, const std::vector<double>& h_u, const std::vector<double>& h_v,
                        const std::vector<double>& h_p, const std::vector<double>& h_uo,
                        const std::vector<double>& h_vo);
    // Upload of a freshly extracted mesh (elements, coordinates, tags): the incidence, the CSR
    // graph, the position table and the mesh checks are built on the device (cfa_mesh_topology,
    // cfa_mesh_check) with the host's results; the graph is copied into mesh for the host users
    // and the cylinder edges are found on the host. valid = the checks passed (the host check
    // then prints what failed). Returns false, with nothing changed on the device side that
    // sync_from_host would not redo, when a row is too long for the device kernel.
    // Upload of a mesh whose CSR graph the host built, with the incidence and position tables
    // formed on the device (cfa_mesh_tables) instead of on the host and uploaded.
    void sync_from_host_device_tables(const Mesh2D& mesh, const std::vector<double>& h_u,
                                      const std::vector<double>& h_v, const std::vector<double>& h_p,
                                      const std::vector<double>& h_uo, const std::vector<double>& h_vo);
    bool sync_from_host_device_topology(Mesh2D& mesh, const std::vector<double>& h_u,
                                        const std::vector<double>& h_v, const std::vector<double>& h_p,
                                        const std::vector<double>& h_uo, const std::vector<double>& h_vo,
                                        MeshQuality& q, bool& valid);
    void copy_solution_to_host(std::vector<double>& h_u, std::vector<double>& h_v, std::vector<double>& h_p,
                               std::vector<double>& h_uo, std::vector<double>& h_vo, std::vector<double>& h_vort,
                               std::vector<double>& h_elem_err);
private:
    cfa_mesh_t view_{};
    DevArray<unsigned char> d_topo_ws_;
    DevArray<double> d_topo_out_;
    DevArray<int> d_topo_nnz_;
    DevArray<double> d_mean_ws_;
    void upload_elements(const Mesh2D& mesh);
    void upload_fields(const Mesh2D& mesh, const std::vector<double>& h_u, const std::vector<double>& h_v,
                       const std::vector<double>& h_p, const std::vector<double>& h_uo,
                       const std::vector<double>& h_vo);
};

// Synchronous remeshes build the CSR graph and check the mesh on the device unless
// CFA_HOST_TOPOLOGY=1 is set (the host path, for comparisons).
bool device_topology_default();

// node -> incident element table of the owner-computes kernels (ascending element order) and
// the CSR position of each (node, element vertex) pair
void build_incidence(const Mesh2D& mesh, std::vector<int>& inc_ptr, std::vector<int>& inc_elem,
                     std::vector<int>& inc_pos);

void compute_error_indicator(DeviceMeshBuffers& g, double inflow_peak = Sim::U_MAX);
void apply_oss_projection_matrix_free(DeviceMeshBuffers& g, const double* in, double* out);
void assemble_oss_projection(DeviceMeshBuffers& g);
void apply_oss_projection(DeviceMeshBuffers& g, const double* in, double* out);
// What the two solves of the scheme-consistent re-projection (u^{n-1} and u^n, independent of
// each other) need to run at the same time: a second stream and solver, the second solve's
// right-hand side, divergence, solution and OSS scratch vectors.
struct ReprojectionAux {
    cfa_stream_t s = nullptr;
    void* ready = nullptr;   // the systems and the hierarchy are built (recorded on the solver stream)
    GpuLinearSolver solver;
    DevArray<double> rhs, div_num, dp, gx, gy;
    ReprojectionAux();
    ~ReprojectionAux();
    ReprojectionAux(const ReprojectionAux&) = delete;
    ReprojectionAux& operator=(const ReprojectionAux&) = delete;
};

int project_scheme_consistent(DeviceMeshBuffers& g, GpuLinearSolver& solver, double dt_proj, double stab,
                              AmgPreconditioner* amg, int& nonconverged,
                              const std::function<void()>& build_amg = nullptr, bool both_levels = true,
                              double* phi_out = nullptr, double tol = 1.0e-8, ReprojectionAux* aux = nullptr);
int project_divergence_free(DeviceMeshBuffers& g, GpuLinearSolver& solver, double dt, AmgPreconditioner* amg,
                            bool setup_amg, double* u, double* v, double* vort_out, double tol = 1.0e-8);
void set_initial_condition(const Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                           std::vector<double>& h_p, double inflow_peak = Sim::U_MAX);
void print_amr_report(const char* tag, const AmrStats& st, const Mesh2D& mesh, const MeshQuality& q,
                      int proj_iters);

// Kernel timings of the last --kernel-timing run or solver workload: name, algorithmic bytes per
// launch, mean duration (us), launches timed.
struct KernelTime {
    std::string name;
    double bytes, us;
    int launches;
};
std::vector<KernelTime>& kernel_timings();

// Mean duration of `launch` over `reps` launches between two HIP events on the solver stream,
// after three untimed launches; microseconds.
template <typename F>
double time_launches(int reps, F launch) {
    const cfa_stream_t s = solver_stream();
    for (int k = 0; k < 3; ++k) launch();
    void *t0 = nullptr, *t1 = nullptr;
    GPU_CHECK(cfa_timer_create(&t0));
    GPU_CHECK(cfa_timer_create(&t1));
    GPU_CHECK(cfa_timer_record(t0, s));
    for (int k = 0; k < reps; ++k) launch();
    GPU_CHECK(cfa_timer_record(t1, s));
    GPU_CHECK(cfa_stream_sync(s));
    float ms = 0.0f;
    GPU_CHECK(cfa_timer_elapsed_ms(t0, t1, &ms));
    GPU_CHECK(cfa_timer_destroy(t0));
    GPU_CHECK(cfa_timer_destroy(t1));
    return 1000.0 * ms / reps;
}

struct RemeshProfile {
    int count = 0;
    double indicator = 0.0;
    double lookahead = 0.0, mark = 0.0, neighbours = 0.0, coarsen = 0.0, refine = 0.0;
    double extract = 0.0, csr = 0.0, validate = 0.0, upload = 0.0, colouring = 0.0;
    double device_topology = 0.0;   // upload, CSR graph and checks built on the device
    double amg_setup = 0.0, projection = 0.0;
    double overlapped = 0.0, join_wait = 0.0, transfer = 0.0;
    void add(const AmrStats& st) {
        ++count;
        lookahead += st.ms_lookahead;
        mark += st.ms_mark;
        neighbours += st.ms_neighbours;
        coarsen += st.ms_coarsen;
        refine += st.ms_refine;
    }
    void report(double total_ms) const;
};

}  // namespace cfa
