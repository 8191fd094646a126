// Time-stepping driver of the command-line program (src/cylinder_fem_amr.cu:119-819): set-up,
// Steps A-D of the BDF2 rotational incremental projection, remeshing (synchronous or
// overlapped with a host worker thread), outputs and the [Timing] report. All device work goes
// through the C ABI of include/cfa_hip.h.
#include <algorithm>
#include <fstream>
#include <future>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "device.hpp"
#include "options.hpp"
#include "outputs.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cfa {

// Figures of the last run, for callers that drive the program in-process (benchmark, tests)
struct RunStats {
    double wall_s = 0, momentum_s = 0, pressure_s = 0, amr_s = 0, amg_setup_s = 0;
    int steps = 0, nodes = 0, elems = 0, nonconverged = 0;
    long long it_u = 0, it_v = 0, it_p = 0;
    double last_cd = 0, last_cl = 0, last_dp = 0;
    double mms[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // sqrt of the six MMS integrals, cylinder/probe max, h_max
    // one row per step: step, time, C_D, C_L, delta_p, nodes, elements, p_front, p_back,
    // C_D_surface, C_L_surface, it_u, it_v, it_p, seconds since the loop started (end of the step)
    static constexpr int HIST_COLS = 15;
    std::vector<double> hist;
};
static RunStats g_stats;

namespace {

// MMS values on the host for the initial condition and the final error report
struct MmsHost { double u, v, p; };
MmsHost mms_exact_host(double x, double y, double t) {
    const double c0 = 2.4390243902439024 * x - 5.3658536585365854;
    const double c1 = 3.1415926535897932 * c0;
    const double c2 = std::cos(c1);
    const double c3 = 15.324842212633138 * x - 33.714652867792903;
    const double c4 = std::cos(c3);
    const double c5 = c2 - 0.25 * c4;
    const double c6 = 7.6624211063165689;
    const double c7 = c6 * y;
    const double c8 = std::sin(c7);
    const double c9 = std::cos(c7);
    const double c10 = -3.8312105531582844 * c8 + 7.6624211063165689 * c9;
    const double c11 = c10 * c5;
    const double c12 = 31.415926535897932 * t;
    const double c13 = std::sin(c12);
    const double c14 = 0.5 * c13 + 1.0;
    const double c15 = 0.13050705333535418 * c14;
    const double c16 = c8 + 0.5 * c9;
    const double c17 = std::sin(c1);
    const double c18 = std::sin(c3);
    const double c19 = -c17 * c6 + 3.8312105531582844 * c18;
    const double c20 = c16 * c19;
    const double c21 = 0.25 * c13 + 0.5;
    const double c22 = -3.1415926535897932 * c0;
    const double c23 = c21 * std::sin(c22);
    MmsHost r;
    r.u = c11 * c15;
    r.v = -c15 * c20;
    r.p = c23 * c9;
    return r;
}

// What an overlapped remesh hands back at its join. The worker thread writes the members, so
// `job` is declared last: it is destroyed first, and ~future waits for the worker before any
// member it writes goes away (an early return from the time loop with a remesh in flight).
struct PendingRemesh {
    int join_step = -1;
    AmrStats st;
    MeshQuality q;
    Mesh2D mesh;
    std::vector<double> u, v, p, uo, vo;
    std::vector<int> old_vids;
    double extract_ms = 0.0, csr_ms = 0.0, validate_ms = 0.0;
    MulticolourIlu0::Pattern ilu;   // prepared by the worker when the momentum ILU is on
    bool has_ilu = false;
    std::future<bool> job;
    ~PendingRemesh() {
        if (job.valid()) job.wait();
    }
};

// State file at the end of a step: int32 magic 0x43464153, n, ne, step; float64 dt; then x, y
// (float64 [n]), bc_tag (int32 [n]), elements (int32 [4 ne]: v0, v1, v2, gen), and the fields
// u, v, p, u_old, v_old, dp (float64 [n] each): what a fixed-mesh continuation from that step needs.
void write_state(const std::string& path, const Mesh2D& mesh, const DeviceMeshBuffers& g, int step, double dt) {
    const int n = mesh.num_nodes, ne = mesh.num_elems;
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) {
        std::fprintf(stderr, "[Output] cannot open %s\n", path.c_str());
        std::exit(EXIT_FAILURE);
    }
    const int hdr[4] = {0x43464153, n, ne, step};
    f.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    f.write(reinterpret_cast<const char*>(&dt), sizeof(dt));
    f.write(reinterpret_cast<const char*>(mesh.x.data()), n * sizeof(double));
    f.write(reinterpret_cast<const char*>(mesh.y.data()), n * sizeof(double));
    f.write(reinterpret_cast<const char*>(mesh.bc_tag.data()), n * sizeof(int));
    std::vector<int> ev(4 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {
        for (int k = 0; k < 3; ++k) ev[4 * (size_t)e + k] = mesh.elems[e].v[k];
        ev[4 * (size_t)e + 3] = mesh.elems[e].gen;
    }
    f.write(reinterpret_cast<const char*>(ev.data()), ev.size() * sizeof(int));
    for (const DevArray<double>* a : {&g.d_u, &g.d_v, &g.d_p, &g.d_u_old, &g.d_v_old, &g.d_dp}) {
        const std::vector<double> h = a->to_host(n);
        f.write(reinterpret_cast<const char*>(h.data()), n * sizeof(double));
    }
    std::printf("[Dump] state at the end of step %d (%d nodes, %d elements) -> %s\n", step, n, ne, path.c_str());
}

void dump_device(std::ofstream& f, const void* d, size_t bytes) {
    std::vector<char> h(bytes);
    if (bytes) GPU_CHECK(cfa_memcpy(h.data(), d, bytes, 1, solver_stream()));
    f.write(h.data(), (std::streamsize)bytes);
}

}  // namespace

int run_driver(int argc, char** argv) {
    RunOptions opt;
    if (!parse_options(argc, argv, opt)) return 0;
    g_stats = RunStats();
    kernel_timings().clear();
    GPU_CHECK(cfa_init(0));
    std::cout.unsetf(std::ios::floatfield);   // a previous in-process run may have left fixed format
    std::cout << std::setprecision(6);
    if (opt.outlet_dirichlet)   // src/cylinder_fem_amr.cu:122-127
        std::cout << "[Setup] outlet: inflow parabola (Dirichlet), pressure increment pinned at one outlet node\n";
    const double dt = opt.dt;
    opt.amr.lookahead_time = opt.lookahead * opt.amr_interval * dt;
    const double rot_coeff = opt.rotational ? Sim::RHO * Sim::NU : 0.0;
    const double force_scale = -2.0 / (Sim::U_MEAN * Sim::U_MEAN * 2.0 * Sim::R_CYL);
    const double surface_scale = -force_scale;
    int host_threads = 1;
#ifdef _OPENMP
    if (!std::getenv("OMP_NUM_THREADS")) omp_set_num_threads(std::max(1, std::min(8, omp_get_num_procs() / 2)));
    host_threads = omp_get_max_threads();
#endif
    std::cout << "========================================================================\n"
              << "  HIP (gfx950) 2D FEM + AMR Incompressible Navier-Stokes Cylinder Flow Solver\n"
              << "  Domain: [0, " << Sim::L_CHANNEL << "] x [0, " << Sim::H_CHANNEL << "] m | Cylinder: ("
              << Sim::X_CYL << ", " << Sim::Y_CYL << "), R = " << Sim::R_CYL << " m\n"
              << "  Re = " << (Sim::U_MEAN * 2.0 * Sim::R_CYL / Sim::NU) << " | dt = " << dt
              << " s | Steps = " << opt.num_steps << " | AMR Interval = " << opt.amr_interval
              << " (Max Level = " << opt.amr.max_level << ")\n"
              << "  AMR: NVB refine+coarsen | tol vorticity " << opt.amr.tol_vort << " m/s, interpolation "
              << opt.amr.tol_interp << " m/s | coarsen below " << opt.amr.coarsen_safety << " x tol | look-ahead "
              << opt.amr.lookahead_time << " s | buffer " << opt.amr.buffer_layers << " | element budget "
              << opt.amr.max_elems << " | remesh "
              << (opt.amr_overlap > 0 ? "overlapped by " + std::to_string(opt.amr_overlap) + " steps" : "synchronous")
              << "\n"
              << "  Time scheme: " << (opt.bdf2 ? "BDF2" : "backward Euler") << " | pressure update: "
              << (opt.rotational ? "rotational" : "standard") << " incremental projection | host threads "
              << host_threads << "\n"
              << "  Solvers: momentum BiCGSTAB + " << (opt.momentum_ilu ? "multicolour ILU(0)" : "Jacobi")
              << (opt.mixed ? " (FP32 factors)" : "") << " | pressure CG + " << (opt.amg ? "SA-AMG" : "Jacobi") << (opt.mixed && opt.amg ? " (FP32 cycle)" : "")
              << " | tol " << opt.mom_tol << " / " << opt.ppe_tol << " / " << opt.proj_tol << "\n"
              << "========================================================================\n";

    // 1. base mesh, initial condition, uniform level / AMR floor (src/cylinder_fem_amr.cu:168-207)
    Mesh2D mesh;
    generate_cylinder_mesh(mesh, 8, 7, 28);
    MeshQuality q0 = finalize_mesh(mesh);
    std::cout << std::fixed << std::setprecision(4) << "[Init] Base Unstructured Mesh: " << mesh.num_nodes
              << " nodes, " << mesh.num_elems << " triangles, " << mesh.nnz << " CSR nonzeros, min angle "
              << std::setprecision(2) << q0.min_angle_deg << " deg.\n"
              << std::setprecision(4);
    const double inflow0 = opt.case_2d3 ? 0.0 : Sim::U_MAX;
    std::vector<double> h_u, h_v, h_p, h_uo, h_vo, h_vort(mesh.num_nodes, 0.0), h_elem_err(2 * mesh.num_elems, 0.0);
    set_initial_condition(mesh, h_u, h_v, h_p, inflow0);
    h_uo = h_u;
    h_vo = h_v;
    AdaptiveMesh amr(mesh, h_u, h_v, h_p);
    amr.extract(mesh, h_u, h_v, h_p, h_uo, h_vo);
    finalize_mesh(mesh);
    if (opt.uniform_level > 0 || opt.amr.min_level > 0) {
        AmrStats st = amr.refine_uniform(std::max(opt.uniform_level, opt.amr.min_level));
        amr.extract(mesh, h_u, h_v, h_p, h_uo, h_vo);
        MeshQuality q = finalize_mesh(mesh);
        set_initial_condition(mesh, h_u, h_v, h_p, inflow0);
        h_uo = h_u;
        h_vo = h_v;
        amr.scatter_solution(h_u, h_v, h_p, h_uo, h_vo);
        print_amr_report("uniform", st, mesh, q, -1);
    }
    if (opt.mms)
        for (int i = 0; i < mesh.num_nodes; ++i) {
            const MmsHost r = mms_exact_host(mesh.x[i], mesh.y[i], 0.0);
            h_u[i] = h_uo[i] = r.u;
            h_v[i] = h_vo[i] = r.v;
            h_p[i] = r.p;
        }
    DeviceMeshBuffers g;
    g.outlet_dirichlet = opt.outlet_dirichlet;
    g.sync_from_host(mesh, h_u, h_v, h_p, h_uo, h_vo);
    GpuLinearSolver solver;
    ReprojectionAux reproj_aux;   // the second re-projection solve's stream, solver and vectors
    if (opt.init_amr > 0)
        for (int cycle = 0; cycle < opt.init_amr; ++cycle) {
            compute_error_indicator(g, inflow0);
            g.copy_solution_to_host(h_u, h_v, h_p, h_uo, h_vo, h_vort, h_elem_err);
            amr.scatter_solution(h_u, h_v, h_p, h_uo, h_vo);
            amr.inlet_umax = inflow0;
            AmrStats st = amr.adapt(h_elem_err, opt.amr);
            amr.extract(mesh, h_u, h_v, h_p, h_uo, h_vo);
            MeshQuality q = finalize_mesh(mesh);
            set_initial_condition(mesh, h_u, h_v, h_p, inflow0);
            h_uo = h_u;
            h_vo = h_v;
            amr.scatter_solution(h_u, h_v, h_p, h_uo, h_vo);
            g.sync_from_host(mesh, h_u, h_v, h_p, h_uo, h_vo);
            print_amr_report("init", st, mesh, q, -1);
        }

    std::ofstream csv;
    if (opt.write_csv) {
        csv.open(opt.prefix + "_forces.csv");
        if (!csv.is_open()) {
            std::fprintf(stderr, "[Output] cannot open %s_forces.csv\n", opt.prefix.c_str());
            return EXIT_FAILURE;
        }
        csv << "step,time,C_D,C_L,delta_p,nodes,elements,p_front,p_back,C_D_surface,C_L_surface\n"
            << std::setprecision(10);
        csv.unsetf(std::ios::floatfield);
    }
    ForceHistory history;
    int nonconverged_solves = 0;
    auto note_nonconverged = [&nonconverged_solves](const GpuLinearSolver& sv, int at_step, const char* what, int iters) {
        if (sv.converged) return;
        if (++nonconverged_solves <= 20)
            std::fprintf(stderr, "[Solver] step %d: %s solve not converged after %d iterations (|r|/|r0| = %.2e)\n",
                         at_step, what, iters, sv.last_rel_residual);
    };
    const bool use_amg = opt.amg;
    AmgPreconditioner amg;
    amg.fp32 = opt.mixed;
    amg.keep_p0_host = opt.amg_remap > 0;   // remap() is the only reader of the host prolongator
    bool amg_dirty = true, phi_pending = false;
    int amg_built_step = 0, amg_remaps_in_row = 0;
    PendingRemesh pend;
    double amg_setup_ms_total = 0.0, amr_ms_total = 0.0, momentum_ms_total = 0.0, pressure_ms_total = 0.0;
    RemeshProfile remesh_prof;
    MulticolourIlu0 mom_ilu;
    mom_ilu.keep_fp32 = opt.mixed;
    GpuBicgstab2 mom_solver;
    mom_solver.precond_fp32 = opt.mixed;
    if (opt.momentum_ilu) mom_ilu.set_pattern(mesh.num_nodes, mesh.csr_row_ptr, mesh.csr_col_ind, g.d_row_ptr, g.d_col_ind);
    const cfa_stream_t s = solver_stream();
    const auto wall_start = Clock::now();

    // 2. time loop (src/cylinder_fem_amr.cu:277-726)
    for (int step = 1; step <= opt.num_steps; ++step) {
        const int n_nodes = g.n_nodes, n_elems = g.n_elems, nnz = g.nnz;
        const cfa_mesh_t* m = g.view();
        const bool use_bdf2 = opt.bdf2 && step > 1;
        const double dt_proj = use_bdf2 ? (2.0 / 3.0) * dt : dt;
        const double a0 = use_bdf2 ? 1.5 / dt : 1.0 / dt;
        const double a1 = use_bdf2 ? 2.0 / dt : 1.0 / dt;
        const double a2 = use_bdf2 ? -0.5 / dt : 0.0;
        const double c1 = use_bdf2 ? 2.0 : 1.0;
        const double c2 = use_bdf2 ? -1.0 : 0.0;
        dev_copy(g.d_u_prev, g.d_u_old, n_nodes);
        dev_copy(g.d_v_prev, g.d_v_old, n_nodes);
        dev_copy(g.d_u_old, g.d_u, n_nodes);
        dev_copy(g.d_v_old, g.d_v, n_nodes);

        // Step A: momentum predictor
        auto t_phase = Clock::now();
        if (opt.mms) GPU_CHECK(cfa_mms_forcing(n_nodes, g.d_x, g.d_y, step * dt, g.d_grad_dp_x, g.d_grad_dp_y, s));
        GPU_CHECK(cfa_assemble_momentum(m, g.d_u_old, g.d_v_old, g.d_u_prev, g.d_v_prev, g.d_p, Sim::NU, Sim::RHO, a0,
                                        a1, a2, c1, c2, opt.supg_tau_time > 0.0 ? opt.supg_tau_time : dt,
                                        opt.skew_convection ? 0.5 : 0.0, opt.mms ? g.d_grad_dp_x.ptr : nullptr,
                                        opt.mms ? g.d_grad_dp_y.ptr : nullptr, inflow_peak(step * dt, opt.case_2d3),
                                        opt.mms ? step * dt : -1.0, g.d_A_val, g.d_rhs_u, g.d_rhs_v, g.d_mass_lumped,
                                        g.d_u, g.d_v, s));
        const bool dump_now = (step == opt.dump_systems);
        std::ofstream dump;
        if (dump_now) {   // int32 n, nnz; rp, ci; val, rhs_u, rhs_v, u0, v0, x, y; bc_tag; then u, v
            dump.open(opt.prefix + "_momentum.bin", std::ios::binary);
            const int hdr[2] = {n_nodes, nnz};
            dump.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
            dump_device(dump, g.d_row_ptr, (n_nodes + 1) * sizeof(int));
            dump_device(dump, g.d_col_ind, nnz * sizeof(int));
            dump_device(dump, g.d_A_val, nnz * sizeof(double));
            for (const double* d : {g.d_rhs_u.ptr, g.d_rhs_v.ptr, g.d_u.ptr, g.d_v.ptr, g.d_x.ptr, g.d_y.ptr})
                dump_device(dump, d, n_nodes * sizeof(double));
            dump_device(dump, g.d_bc_tag, n_nodes * sizeof(int));
        }
        int it_u, it_v;
        if (opt.momentum_ilu) {
            mom_ilu.factor(g.d_A_val);
            mom_solver.solve(n_nodes, g.d_row_ptr, g.d_col_ind, g.d_A_val, g.d_rhs_u, g.d_rhs_v, g.d_u, g.d_v, 150,
                             opt.mom_tol, mom_ilu);
            it_u = mom_solver.iters[0];
            it_v = mom_solver.iters[1];
            for (int c = 0; c < 2; ++c) {
                if (mom_solver.converged[c]) continue;
                if (++nonconverged_solves <= 20)
                    std::fprintf(stderr, "[Solver] step %d: momentum %c solve not converged after %d iterations (|r|/|r0| = %.2e)\n",
                                 step, c ? 'v' : 'u', mom_solver.iters[c], mom_solver.last_rel_residual[c]);
            }
        } else {
            it_u = solver.solve_bicgstab(n_nodes, g.d_row_ptr, g.d_col_ind, g.d_A_val, g.d_rhs_u, g.d_u, 150, opt.mom_tol);
            note_nonconverged(solver, step, "momentum u", it_u);
            it_v = solver.solve_bicgstab(n_nodes, g.d_row_ptr, g.d_col_ind, g.d_A_val, g.d_rhs_v, g.d_v, 150, opt.mom_tol);
            note_nonconverged(solver, step, "momentum v", it_v);
        }
        if (dump_now) {
            dump_device(dump, g.d_u, n_nodes * sizeof(double));
            dump_device(dump, g.d_v, n_nodes * sizeof(double));
            dump.close();
            std::printf("[Dump] momentum system of step %d (%d nodes, nnz %d, iterations %d/%d) -> %s_momentum.bin\n",
                        step, n_nodes, nnz, it_u, it_v, opt.prefix.c_str());
        }
        dev_sync();
        momentum_ms_total += ms_since(t_phase);
        t_phase = Clock::now();

        // Step B: pressure increment
        const bool oss = opt.pressure_stab > 0.0;
        const bool shift_now = phi_pending && oss;
        if (shift_now) GPU_CHECK(cfa_axpy(n_nodes, 1.0, g.d_phi, g.d_p, s));
        phi_pending = false;
        if (oss) {
            GPU_CHECK(cfa_oss_weights(m, g.d_u, g.d_v, dt_proj, Sim::RHO, Sim::NU, opt.pressure_stab, g.d_elem_s, s));
            GPU_CHECK(cfa_scalar_gradient(m, g.d_p, g.d_elem_s, g.d_grad_dp_x, g.d_grad_dp_y, s));
        }
        GPU_CHECK(cfa_assemble_pressure(m, g.d_u, g.d_v, g.d_p, g.d_grad_dp_x, g.d_grad_dp_y, g.d_mass_lumped,
                                        g.d_elem_s, dt_proj, Sim::RHO, opt.pressure_stab, g.d_Kp_val, g.d_rhs_p,
                                        g.d_div_num, g.d_dp, oss ? g.d_oss_val.ptr : nullptr, g.d_inv_m, s));
        g.pressure_outlet_bc(g.d_mass_lumped, g.d_rhs_p, s);
        if (shift_now) GPU_CHECK(cfa_axpy(n_nodes, -1.0, g.d_phi, g.d_p, s));
        if (dump_now) {   // int32 n, nnz; rp, ci; K_p values, rhs, initial guess
            dump.open(opt.prefix + "_pressure.bin", std::ios::binary);
            const int hdr[2] = {n_nodes, nnz};
            dump.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
            dump_device(dump, g.d_row_ptr, (n_nodes + 1) * sizeof(int));
            dump_device(dump, g.d_col_ind, nnz * sizeof(int));
            dump_device(dump, g.d_Kp_val, nnz * sizeof(double));
            dump_device(dump, g.d_rhs_p, n_nodes * sizeof(double));
            dump_device(dump, g.d_dp, n_nodes * sizeof(double));
            dump.close();
        }
        if (use_amg && (amg_dirty || step - amg_built_step >= opt.amg_rebuild)) {
            amg.setup(n_nodes, nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
            amg_dirty = false;
            amg_built_step = step;
            amg_setup_ms_total += amg.setup_ms;
        }
        const cfa_oss_op_t op = g.oss_op();
        int it_p;
        if (use_amg) {
            it_p = solver.solve_pcg_fused(n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000,
                                          opt.ppe_tol, oss ? &op : nullptr, &amg);
        } else {
            std::function<void(const double*, double*)> oss_fn = [&g](const double* in, double* out) {
                apply_oss_projection(g, in, out);
            };
            it_p = solver.solve_pcg(n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000,
                                    opt.ppe_tol, oss ? &oss_fn : nullptr, nullptr);
        }
        note_nonconverged(solver, step, "pressure", it_p);
        if (opt.rotational && opt.rotational_from_dp)
            GPU_CHECK(cfa_stiffness_apply(m, g.d_dp, -dt_proj / Sim::RHO, g.d_div_num, s));
        dev_sync();
        pressure_ms_total += ms_since(t_phase);

        // Step C: correction
        GPU_CHECK(cfa_correction(m, dt_proj, Sim::RHO, 1.0, rot_coeff, g.d_div_num, g.d_mass_lumped, g.d_dp, g.d_u,
                                 g.d_v, g.d_p, g.d_vort, g.d_grad_dp_x, g.d_grad_dp_y, g.d_vort_num, s));

        // Step D: forces and probes
        const int probe[4] = {g.probe_front.a, g.probe_front.b, g.probe_back.a, g.probe_back.b};
        GPU_CHECK(cfa_forces(m, g.d_cyl_edges, g.n_cyl_edges, g.d_layer, g.n_layer, g.d_u, g.d_v, g.d_u_old, g.d_v_old,
                             g.d_u_prev, g.d_v_prev, g.d_p, a0, -a1, -a2, Sim::NU, Sim::RHO, probe, g.probe_front.w,
                             g.probe_back.w, g.d_diag, g.d_ws, s));
        double diag[6];
        g.d_diag.download(diag, 6);
        const double t_now = step * dt;
        const double C_D = force_scale * diag[0], C_L = force_scale * diag[1], delta_p = diag[2] - diag[3];
        if (!std::isfinite(C_D) || !std::isfinite(C_L) || !std::isfinite(delta_p)) {
            std::fprintf(stderr, "[Step %d] non-finite solution detected, aborting\n", step);
            return EXIT_FAILURE;
        }
        g_stats.it_u += it_u;
        g_stats.it_v += it_v;
        g_stats.it_p += it_p;
        g_stats.last_cd = C_D;
        g_stats.last_cl = C_L;
        g_stats.last_dp = delta_p;
        const size_t hist_row = g_stats.hist.size();
        for (double val : {(double)step, t_now, C_D, C_L, delta_p, (double)n_nodes, (double)n_elems, diag[2], diag[3],
                           surface_scale * diag[4], surface_scale * diag[5], (double)it_u, (double)it_v, (double)it_p,
                           0.0})
            g_stats.hist.push_back(val);
        if (csv.is_open())
            csv << step << "," << t_now << "," << C_D << "," << C_L << "," << delta_p << "," << n_nodes << ","
                << n_elems << "," << diag[2] << "," << diag[3] << "," << surface_scale * diag[4] << ","
                << surface_scale * diag[5] << "\n";
        if (!opt.quiet && (step % opt.print_interval == 0 || step == 1 || step == opt.num_steps))
            std::cout << std::fixed << std::setprecision(4) << "[Step " << std::setw(5) << step << " | t = " << t_now
                      << " s] Nodes: " << std::setw(6) << n_nodes << " | Elems: " << std::setw(6) << n_elems
                      << " | Iters(u,v,p): (" << it_u << "," << it_v << "," << it_p << ") | C_D = " << C_D
                      << " | C_L = " << C_L << " | dP = " << delta_p << std::endl;

        history.add(t_now, C_D, C_L, delta_p);
        if (step == opt.dump_state) write_state(opt.prefix + "_state.bin", mesh, g, step, dt);
        // Step E: VTK snapshots (src/cylinder_fem_amr.cu:540-547)
        if (opt.write_csv && opt.vtk_interval > 0 && step % opt.vtk_interval == 0 && step < opt.num_steps) {
            compute_error_indicator(g);
            g.copy_solution_to_host(h_u, h_v, h_p, h_uo, h_vo, h_vort, h_elem_err);
            std::ostringstream name;
            name << opt.prefix << "_" << std::setw(6) << std::setfill('0') << step << ".vtk";
            write_vtk(name.str(), mesh, h_u, h_v, h_p, h_vort, h_elem_err, opt.amr);
        }

        // Step F: remeshing (src/cylinder_fem_amr.cu:549-725)
        const bool overlap = opt.amr_overlap > 0;
        const bool remesh_due = opt.amr_interval > 0 && step % opt.amr_interval == 0 && step < opt.num_steps &&
                                !(overlap && step + opt.amr_overlap >= opt.num_steps);
        auto t0 = Clock::now();
        auto tp = t0;
        auto lap = [&tp]() { const double ms = ms_since(tp); tp = Clock::now(); return ms; };
        const bool amg_remap_now = use_amg && opt.remesh_projection >= 3 && opt.amg_remap > 0 &&
                                   amg_remaps_in_row < opt.amg_remap;
        auto reproject = [&](const std::function<void()>& build_amg) -> int {
            int proj_iters = 0;
            const bool scheme_proj = opt.remesh_projection >= 3;
            if (scheme_proj) {
                const int remaps_before = amg.remaps;
                proj_iters = project_scheme_consistent(g, solver, opt.bdf2 ? (2.0 / 3.0) * dt : dt, opt.pressure_stab,
                                                       use_amg ? &amg : nullptr, nonconverged_solves, build_amg,
                                                       opt.remesh_projection == 4,
                                                       opt.reproject_shift ? g.d_phi.ptr : nullptr, opt.proj_tol,
                                                       &reproj_aux);
                amg_remaps_in_row = amg.remaps > remaps_before ? amg_remaps_in_row + 1 : 0;
                phi_pending = opt.reproject_shift;
            }
            if (!scheme_proj && opt.remesh_projection >= 2) {
                proj_iters += project_divergence_free(g, solver, dt, use_amg ? &amg : nullptr, true, g.d_u_old,
                                                      g.d_v_old, g.d_oss_gx, opt.proj_tol);
                nonconverged_solves += solver.converged ? 0 : 1;
            }
            if (!scheme_proj && opt.remesh_projection >= 1)
                proj_iters += project_divergence_free(g, solver, dt, use_amg ? &amg : nullptr,
                                                      opt.remesh_projection == 1, g.d_u, g.d_v, g.d_vort, opt.proj_tol);
            amg_dirty = true;
            if (scheme_proj && use_amg) {
                amg_dirty = false;
                amg_built_step = step;
            }
            if (!scheme_proj) nonconverged_solves += solver.converged ? 0 : 1;
            return proj_iters;
        };
        auto finish_remesh = [&](const AmrStats& st, const MeshQuality& q, const std::vector<int>& old_vids,
                                 const std::vector<double>& old_x, const std::vector<double>& old_y,
                                 bool uploaded) {
            if (!uploaded) {   // the graph from the host; the tables built on the device by default
                if (device_topology_default()) g.sync_from_host_device_tables(mesh, h_u, h_v, h_p, h_uo, h_vo);
                else g.sync_from_host(mesh, h_u, h_v, h_p, h_uo, h_vo);
                dev_sync();
                remesh_prof.upload += lap();
            }
            if (opt.momentum_ilu) {
                if (pend.has_ilu) {   // colouring done by the remesh worker, only the upload here
                    mom_ilu.set_pattern(pend.ilu, g.d_row_ptr, g.d_col_ind);
                    pend.ilu = MulticolourIlu0::Pattern();
                    pend.has_ilu = false;
                } else {
                    mom_ilu.set_pattern(mesh.num_nodes, mesh.csr_row_ptr, mesh.csr_col_ind, g.d_row_ptr, g.d_col_ind);
                }
            }
            remesh_prof.colouring += lap();
            std::function<void()> build_amg;
            if (amg_remap_now)
                build_amg = [&]() {
                    std::vector<int> slot_to_old(amr.num_vertex_slots(), -1), new_to_old(mesh.num_nodes);
                    for (int j = 0; j < (int)old_vids.size(); ++j)
                        if (old_vids[j] < (int)slot_to_old.size()) slot_to_old[old_vids[j]] = j;
                    for (int i = 0; i < mesh.num_nodes; ++i) {
                        const int j = slot_to_old[amr.node_vertices()[i]];
                        new_to_old[i] = (j >= 0 && old_x[j] == mesh.x[i] && old_y[j] == mesh.y[i]) ? j : -1;
                    }
                    amg.remap(g.n_nodes, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, mesh.csr_row_ptr.data(),
                              mesh.csr_col_ind.data(), new_to_old);
                };
            const int proj_iters = reproject(build_amg);
            const double proj_ms = lap();
            const double proj_amg_ms = (use_amg && opt.remesh_projection > 0) ? amg.setup_ms : 0.0;
            remesh_prof.amg_setup += proj_amg_ms;
            remesh_prof.projection += proj_ms - proj_amg_ms;
            amr_ms_total += ms_since(t0);
            print_amr_report("remesh", st, mesh, q, proj_iters);
        };

        if (opt.reproject_test > 0 && step % opt.reproject_test == 0 && step < opt.num_steps) {
            const int n = g.n_nodes;
            std::vector<double> u0 = g.d_u.to_host(n), v0 = g.d_v.to_host(n);
            reproject(nullptr);
            std::vector<double> u1 = g.d_u.to_host(n), v1 = g.d_v.to_host(n);
            double du = 0.0, um = 0.0;
            for (int i = 0; i < n; ++i) {
                du = std::max(du, std::hypot(u1[i] - u0[i], v1[i] - v0[i]));
                um = std::max(um, std::hypot(u0[i], v0[i]));
            }
            std::printf("[Reproject-test] step %d: max|u_new - u| = %.3e (max|u| = %.3e)\n", step, du, um);
        }

        if (remesh_due) {
            compute_error_indicator(g, inflow_peak(step * dt, opt.case_2d3));
            amr.inlet_umax = inflow_peak(step * dt, opt.case_2d3);
            g.copy_solution_to_host(h_u, h_v, h_p, h_uo, h_vo, h_vort, h_elem_err);
            amr.scatter_solution(h_u, h_v, h_p, h_uo, h_vo);
            remesh_prof.indicator += lap();
            AmrParams prm_now = opt.amr;
            if (opt.case_2d3 && opt.relative_tol) {
                const double scale = std::max(0.1, std::fabs(inflow_peak(step * dt, true)) / Sim::U_MAX);
                prm_now.tol_vort *= scale;
                prm_now.tol_interp *= scale;
            }
            if (!overlap) {
                AmrStats st = amr.adapt(h_elem_err, prm_now);
                remesh_prof.add(st);
                std::vector<int> old_vids;
                std::vector<double> old_x, old_y;
                if (amg_remap_now) { old_vids = amr.node_vertices(); old_x = mesh.x; old_y = mesh.y; }
                lap();
                amr.extract(mesh, h_u, h_v, h_p, h_uo, h_vo);
                remesh_prof.extract += lap();
                MeshQuality q;
                bool valid = true;
                // the graph, the position table and the checks on the device, the host path when
                // a row is too long for it (or CFA_HOST_TOPOLOGY=1)
                const bool on_device = device_topology_default() &&
                                       g.sync_from_host_device_topology(mesh, h_u, h_v, h_p, h_uo, h_vo, q, valid);
                if (on_device) {
                    dev_sync();
                    remesh_prof.device_topology += lap();
                } else {
                    build_csr_and_boundary_topology(mesh);
                    remesh_prof.csr += lap();
                    valid = validate_mesh(mesh, q);
                    remesh_prof.validate += lap();
                }
                if (!valid) {
                    std::fprintf(stderr, "[AMR] mesh validation failed, aborting\n");
                    return EXIT_FAILURE;
                }
                finish_remesh(st, q, old_vids, old_x, old_y, on_device);
            } else {
                prm_now.lookahead_time += opt.amr_overlap * dt;
                pend.old_vids = amr.node_vertices();
                pend.join_step = step + opt.amr_overlap;
                // the worker owns the forest and pend until the join; the forest's passes run on
                // its own stream, which does not synchronise with the solver's
                const bool want_ilu = opt.momentum_ilu;
                pend.job = std::async(std::launch::async, [&amr, &pend, prm_now, want_ilu, err = h_elem_err]() {
                    auto tw = Clock::now();
                    pend.st = amr.adapt(err, prm_now);
                    tw = Clock::now();
                    amr.extract(pend.mesh, pend.u, pend.v, pend.p, pend.uo, pend.vo);
                    pend.extract_ms = ms_since(tw);
                    tw = Clock::now();
                    build_csr_and_boundary_topology(pend.mesh);
                    pend.csr_ms = ms_since(tw);
                    tw = Clock::now();
                    const bool ok = validate_mesh(pend.mesh, pend.q);
                    pend.validate_ms = ms_since(tw);
                    if (ok && want_ilu) {
                        pend.ilu = MulticolourIlu0::build_pattern(pend.mesh.num_nodes, pend.mesh.csr_row_ptr,
                                                                  pend.mesh.csr_col_ind);
                        pend.has_ilu = true;
                    }
                    return ok;
                });
                amr_ms_total += ms_since(t0);
            }
        }
        if (overlap && step == pend.join_step) {
            tp = t0 = Clock::now();
            const bool ok = pend.job.get();
            remesh_prof.join_wait += lap();
            if (!ok) {
                std::fprintf(stderr, "[AMR] mesh validation failed, aborting\n");
                return EXIT_FAILURE;
            }
            remesh_prof.add(pend.st);
            remesh_prof.extract += pend.extract_ms;
            remesh_prof.csr += pend.csr_ms;
            remesh_prof.validate += pend.validate_ms;
            remesh_prof.overlapped += pend.st.ms_lookahead + pend.st.ms_mark + pend.st.ms_neighbours +
                                      pend.st.ms_coarsen + pend.st.ms_refine + pend.extract_ms + pend.csr_ms +
                                      pend.validate_ms + (pend.has_ilu ? pend.ilu.build_ms : 0.0);
            g.copy_solution_to_host(h_u, h_v, h_p, h_uo, h_vo, h_vort, h_elem_err);
            amr.inlet_umax = inflow_peak(step * dt, opt.case_2d3);
            std::vector<double> old_x, old_y;
            if (amg_remap_now) { old_x = mesh.x; old_y = mesh.y; }
            amr.retransfer(pend.old_vids, h_u, h_v, h_p, h_uo, h_vo, pend.u, pend.v, pend.p, pend.uo, pend.vo);
            mesh = std::move(pend.mesh);
            pend.mesh = Mesh2D();
            h_u.swap(pend.u);
            h_v.swap(pend.v);
            h_p.swap(pend.p);
            h_uo.swap(pend.uo);
            h_vo.swap(pend.vo);
            remesh_prof.transfer += lap();
            pend.join_step = -1;
            finish_remesh(pend.st, pend.q, pend.old_vids, old_x, old_y, false);
        }
        g_stats.hist[hist_row + RunStats::HIST_COLS - 1] = ms_since(wall_start) / 1000.0;
    }
    if (csv.is_open()) csv.close();
    dev_sync();
    const double wall_s = ms_since(wall_start) / 1000.0;

    if (opt.write_csv) {   // final solution and mesh (src/cylinder_fem_amr.cu:731-735)
        compute_error_indicator(g);
        g.copy_solution_to_host(h_u, h_v, h_p, h_uo, h_vo, h_vort, h_elem_err);
        write_vtk(opt.prefix + "_final.vtk", mesh, h_u, h_v, h_p, h_vort, h_elem_err, opt.amr);
    }
    if (!opt.mms) {   // DFG statistics of the force history (src/cylinder_fem_amr.cu:784-788)
        const std::string stats_csv = opt.write_csv ? opt.prefix + "_stats.csv" : std::string();
        if (opt.case_2d3) history.report_2d3(stats_csv);
        else history.report(opt.stats_window, stats_csv);
    }
    if (opt.mms) {
        double e2[6];
        DevArray<double> d_err;
        d_err.reserve(6, 0);
        GPU_CHECK(cfa_mms_errors(g.view(), g.d_u, g.d_v, g.d_p, opt.num_steps * dt, d_err, g.d_ws, s));
        d_err.download(e2, 6);
        std::vector<double> hp = g.d_p.to_host(g.n_nodes);
        const double t_end = opt.num_steps * dt;
        double cyl_max = 0.0;
        for (int i = 0; i < mesh.num_nodes; ++i) {
            if (mesh.bc_tag[i] != BC_CYLINDER) continue;
            cyl_max = std::max(cyl_max, std::fabs(hp[i] - mms_exact_host(mesh.x[i], mesh.y[i], t_end).p));
        }
        auto probe_err = [&](const PressureProbe& pr, double px, double py) {
            return std::fabs((1.0 - pr.w) * hp[pr.a] + pr.w * hp[pr.b] - mms_exact_host(px, py, t_end).p);
        };
        const double probe_max = std::max(probe_err(g.probe_front, 0.15, 0.20), probe_err(g.probe_back, 0.25, 0.20));
        double h_max = 0.0;
        for (const auto& tri : mesh.elems)
            for (int k = 0; k < 3; ++k) {
                const int a = tri.v[k], b = tri.v[(k + 1) % 3];
                h_max = std::max(h_max, std::hypot(mesh.x[a] - mesh.x[b], mesh.y[a] - mesh.y[b]));
            }
        std::cout << std::scientific << std::setprecision(6) << "[MMS] t = " << t_end << " | nodes " << mesh.num_nodes
                  << " | h_max " << h_max << " | dt " << dt << "\n"
                  << "[MMS] ||u-u_h||_L2 = " << std::sqrt(e2[0]) << " (rel " << std::sqrt(e2[0] / e2[3]) << ")"
                  << " | |u-u_h|_H1 = " << std::sqrt(e2[1]) << " (rel " << std::sqrt(e2[1] / e2[4]) << ")"
                  << " | ||p-p_h||_L2 = " << std::sqrt(e2[2]) << " (rel " << std::sqrt(e2[2] / e2[5]) << ")\n"
                  << "[MMS] max |p-p_h| on cylinder nodes = " << cyl_max << " | at the DFG probes = " << probe_max
                  << "\n"
                  << std::fixed;
        for (int k = 0; k < 6; ++k) g_stats.mms[k] = std::sqrt(e2[k]);
        g_stats.mms[6] = cyl_max;
        g_stats.mms[7] = probe_max;
        g_stats.mms[8] = h_max;
        if (opt.write_csv) {
            std::ofstream f(opt.prefix + "_mms.csv");
            f << "nodes,elements,h_max,dt,t,err_u_L2,err_u_H1,err_p_L2,norm_u_L2,norm_u_H1,norm_p_L2,err_p_cyl_max,"
                 "err_p_probe_max\n"
              << std::setprecision(10);
            f.unsetf(std::ios::floatfield);
            f << mesh.num_nodes << "," << mesh.num_elems << "," << h_max << "," << dt << "," << opt.num_steps * dt;
            for (int k = 0; k < 6; ++k) f << "," << std::sqrt(e2[k]);
            f << "," << cyl_max << "," << probe_max << "\n";
        }
    }
    if (opt.kernel_timing > 0) {
        // The main kernels of the step on the final mesh and state, each launched alone. Bytes
        // count every array a kernel reads or writes once (n rows, nnz entries, ne elements).
        const int R = opt.kernel_timing;
        const double n = g.n_nodes, nnz = g.nnz, ne = g.n_elems;
        DevArray<double> t_in, t_out, t_part;
        t_in.reserve(g.n_nodes, 0);
        t_out.reserve(g.n_nodes, 0);
        t_part.reserve(4096, 0);
        dev_copy(t_in, g.d_dp, g.n_nodes);
        auto add = [](const char* name, double bytes, double us, int launches) {
            kernel_timings().push_back({name, bytes, us, launches});
        };
        if (opt.pressure_stab > 0.0) {
            const cfa_oss_op_t op = g.oss_op();
            add("k_apply_K_B", 28.0 * nnz + 44.0 * n + 4.0, time_launches(R, [&] {
                    GPU_CHECK(cfa_pcg_operator(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, &op, t_in, t_out,
                                               t_part, 1, s));
                }), R);
            add("k_apply_BT_dot", 20.0 * nnz + 48.0 * n + 4.0, time_launches(R, [&] {
                    GPU_CHECK(cfa_pcg_operator(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, &op, t_in, t_out,
                                               t_part, 2, s));
                }), R);
        }
        if (use_amg && !amg.levels.empty()) {
            const double n0 = amg.levels[0]->A.rows, nnz0 = amg.levels[0]->A.nnz;
            add(amg.levels[0]->has_sell ? "level-0 Jacobi sweep (k_sell_op, sliced ELL)"
                                        : "level-0 Jacobi sweep (k_csr_op)",
                12.0 * nnz0 + 36.0 * n0 + 4.0, time_launches(R, [&] {
                    GPU_CHECK(cfa_amg_smooth(amg.view(), 0, t_in, g.d_rhs_p, t_out, s));
                }), R);
        }
        // plain device copy of n doubles: the reference point for the HBM traffic counters
        add("copy (hipMemcpy device to device)", 16.0 * n, time_launches(R, [&] { dev_copy(t_out, t_in, g.n_nodes); }),
            R);
        const double t_end = opt.num_steps * dt;
        const bool bdf = opt.bdf2 && opt.num_steps > 1;
        add("k_momentum_rows", 12.0 * nnz + 92.0 * n + 60.0 * ne + (opt.mms ? 16.0 * n : 0.0), time_launches(R, [&] {
                GPU_CHECK(cfa_assemble_momentum(g.view(), g.d_u_old, g.d_v_old, g.d_u_prev, g.d_v_prev, g.d_p, Sim::NU,
                                                Sim::RHO, bdf ? 1.5 / dt : 1.0 / dt, bdf ? 2.0 / dt : 1.0 / dt,
                                                bdf ? -0.5 / dt : 0.0, bdf ? 2.0 : 1.0, bdf ? -1.0 : 0.0,
                                                opt.supg_tau_time > 0.0 ? opt.supg_tau_time : dt,
                                                opt.skew_convection ? 0.5 : 0.0, opt.mms ? g.d_grad_dp_x.ptr : nullptr,
                                                opt.mms ? g.d_grad_dp_y.ptr : nullptr, inflow_peak(t_end, opt.case_2d3),
                                                opt.mms ? t_end : -1.0, g.d_A_val, g.d_rhs_u, g.d_rhs_v,
                                                g.d_mass_lumped, g.d_u, g.d_v, s));
            }), R);
        std::cout << "[Timing] kernels on the final mesh (" << g.n_nodes << " nodes, nnz " << g.nnz << ", " << R
                  << " launches each):";
        for (const auto& k : kernel_timings())
            std::cout << " " << k.name << " " << std::setprecision(2) << k.us << " us = " << std::setprecision(0)
                      << k.bytes / (k.us * 1.0e3) << " GB/s;";
        std::cout << "\n";
    }
    std::cout << std::fixed << std::setprecision(2) << "[Timing] " << wall_s << " s wall, " << amr_ms_total / 1000.0
              << " s in AMR (incl. transfer) | PPE AMG setup " << amg_setup_ms_total / 1000.0 << " s | AMG remaps "
              << amg.remaps << " | non-converged linear solves: " << nonconverged_solves << "\n";
    std::cout << "[Timing] step phases: momentum " << momentum_ms_total / 1000.0 << " s"
              << (opt.momentum_ilu ? " (ILU colouring " : "");
    if (opt.momentum_ilu) std::cout << mom_ilu.colour_ms / 1000.0 << " s, " << mom_ilu.n_colours << " colours at the end)";
    std::cout << " | pressure " << pressure_ms_total / 1000.0 << " s (assembly, AMG setup, CG)\n";
    remesh_prof.report(amr_ms_total);
    {
        static const char* dn[] = {"marking", "coarsening", "refinement", "leaf lists", "extraction",
                                   "mesh download"};
        std::cout << "[Timing] AMR sub-phases:";
        for (int i = 0; i < AdaptiveMesh::D_COUNT; ++i)
            std::cout << " " << dn[i] << " " << std::setprecision(2) << amr.dbg_ms[i] / 1000.0 << " s;";
        std::cout << "\n";
    }
    if (use_amg) {
        static const char* names[] = {"download", "D^-1 + rho", "aggregation", "P smoothing", "transpose",
                                      "Galerkin product", "upload", "coarse inverse", "remap"};
        std::cout << "[Timing] AMG setup phases (all setups):";
        for (int i = 0; i < AmgPreconditioner::PH_COUNT; ++i)
            std::cout << " " << names[i] << " " << std::setprecision(2) << amg.phase_ms[i] / 1000.0 << " s;";
        std::cout << "\n";
    }
    if (opt.write_csv) std::cout << "[Output] Wrote force history to: " << opt.prefix << "_forces.csv\n";
    g_stats.wall_s = wall_s;
    g_stats.momentum_s = momentum_ms_total / 1000.0;
    g_stats.pressure_s = pressure_ms_total / 1000.0;
    g_stats.amr_s = amr_ms_total / 1000.0;
    g_stats.amg_setup_s = amg_setup_ms_total / 1000.0;
    g_stats.steps = opt.num_steps;
    g_stats.nodes = g.n_nodes;
    g_stats.elems = g.n_elems;
    g_stats.nonconverged = nonconverged_solves;
    return 0;
}

}  // namespace cfa

extern "C" {

// Runs the command-line program in-process (argv[0] is the program name).
int cfa_main(int argc, char** argv) { return cfa::run_driver(argc, argv); }

// Figures of the last cfa_main run: out[16] = wall s, momentum s, pressure s, AMR s, AMG set-up s,
// steps, nodes, elements, non-converged solves, sum it_u, sum it_v, sum it_p, last C_D, C_L, dP.
void cfa_last_run_stats(double* out) {
    const cfa::RunStats& r = cfa::g_stats;
    const double v[15] = {r.wall_s, r.momentum_s, r.pressure_s, r.amr_s, r.amg_setup_s, (double)r.steps,
                          (double)r.nodes, (double)r.elems, (double)r.nonconverged, (double)r.it_u,
                          (double)r.it_v, (double)r.it_p, r.last_cd, r.last_cl, r.last_dp};
    for (int k = 0; k < 15; ++k) out[k] = v[k];
    out[15] = 0.0;
}

// Per-step history of the last run (rows of 15 values, see RunStats::hist); copies at most
// max_rows rows into out (may be NULL) and returns the number of rows recorded.
int cfa_last_run_history(double* out, int max_rows) {
    const auto& h = cfa::g_stats.hist;
    const int rows = (int)(h.size() / cfa::RunStats::HIST_COLS);
    if (out && max_rows > 0) {
        const int k = std::min(rows, max_rows);
        std::copy(h.begin(), h.begin() + (size_t)k * cfa::RunStats::HIST_COLS, out);
    }
    return rows;
}

// --kernel-timing figures of the last run: names (joined by '\n') into names[cap] and rows of
// (algorithmic bytes per launch, mean microseconds, launches) into out; returns the kernel count.
int cfa_last_run_kernels(char* names, int cap, double* out, int max_rows) {
    const auto& ks = cfa::kernel_timings();
    std::string joined;
    for (size_t k = 0; k < ks.size(); ++k) {
        if (k) joined += '\n';
        joined += ks[k].name;
        if ((int)k < max_rows && out) {
            out[3 * k] = ks[k].bytes;
            out[3 * k + 1] = ks[k].us;
            out[3 * k + 2] = ks[k].launches;
        }
    }
    if (names && cap > 0) {
        const size_t len = std::min(joined.size(), (size_t)cap - 1);
        std::copy(joined.begin(), joined.begin() + len, names);
        names[len] = '\0';
    }
    return (int)ks.size();
}

// MMS figures of the last --mms run: out[9] = ||u-u_h||_L2, |u-u_h|_H1, ||p-p_h||_L2, ||u||_L2,
// |u|_H1, ||p||_L2, max |p-p_h| on the cylinder, at the probes, h_max
void cfa_last_run_mms(double* out) {
    for (int k = 0; k < 9; ++k) out[k] = cfa::g_stats.mms[k];
}

}  // extern "C"
