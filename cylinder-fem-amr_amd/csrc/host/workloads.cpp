// Isolated solver workloads on a frozen uniform mesh: the pressure Poisson system with SA-AMG PCG
// and the coupled momentum system with multicolour-ILU(0) BiCGSTAB (FP64 and FP32 factors). The
// systems are assembled by the step's own kernels (src/cfa/fem_kernels.cuh
// k_assemble_pressure_poisson with stab = 0 and apply_bc_pressure; k_assemble_momentum_fem with
// k_apply_bc_momentum) and solved by the step's own solver classes, so timings per right-hand
// side and per Krylov iteration measure exactly the path's solvers at a chosen mesh size.
#include <cmath>
#include <random>

#include "device.hpp"

namespace cfa {
namespace {

Mesh2D uniform_mesh(int levels) {
    Mesh2D m;
    generate_cylinder_mesh(m, 8, 7, 28);
    finalize_mesh(m);
    std::vector<double> z(m.num_nodes, 0.0), u = z, v = z, p = z, uo = z, vo = z;
    AdaptiveMesh amr(m, z, z, z);
    amr.extract(m, u, v, p, uo, vo);
    finalize_mesh(m);
    if (levels > 0) {
        amr.refine_uniform(levels);
        amr.extract(m, u, v, p, uo, vo);
        finalize_mesh(m);
    }
    return m;
}

// phase times (ms) of the last pressure workload's AMG set-up, then the levels built on the device
double g_amg_phases[AmgPreconditioner::PH_COUNT + 3] = {};

// Host copies of the last workload systems (cfa_workload_keep_systems): the matrix, the
// right-hand sides, the initial guesses and the solutions, for checks outside the library.
struct KeptSystem {
    int n = 0, nnz = 0, cols = 0;    // cols: right-hand sides (pressure) or 2 (momentum)
    std::vector<int> rp, ci, tag;
    std::vector<double> val, b, x0, x;   // b, x0: cols x n; x: pressure cols x n, momentum 4 x n
    std::vector<int> iters;              // per solution of x
};
bool g_keep = false;
KeptSystem g_kept[2];   // 0 pressure, 1 momentum

void keep_pattern(KeptSystem& k, const Mesh2D& mesh, const DevArray<double>& val, int nnz) {
    k = KeptSystem();
    k.n = mesh.num_nodes;
    k.nnz = nnz;
    k.rp = mesh.csr_row_ptr;
    k.ci = mesh.csr_col_ind;
    k.tag = mesh.bc_tag;
    k.val = val.to_host(nnz);
}

double host_ms(Clock::time_point t0) {
    dev_sync();
    return ms_since(t0);
}

}  // namespace
}  // namespace cfa

using namespace cfa;

extern "C" {

// Pressure workload: P1 Laplacian of the PPE (stab = 0, outlet rows Dirichlet) on the uniform
// level-`level` mesh, one SA-AMG set-up, then `nrhs` solves from x = 0 to relative residual `tol`
// with right-hand sides drawn from mt19937_64(1000 + k) N(0,1), mean over the free rows removed,
// Dirichlet rows zero. out[12] = n, nnz, ne, mesh s, AMG set-up ms, AMG levels, solves,
// solve ms total, iterations total, ms per solve, us per iteration, max final relative residual.
// With kernel_reps > 0 the SpMV, one smoothing sweep per level and lanes-per-row choice, a 1 GiB
// device copy and, last, the level-0 smoother as the V-cycle runs it are timed afterwards
// (kernel_timings()).
int cfa_workload_pressure(int level, int nrhs, double tol, int kernel_reps, double* out) {
    if (level < 0 || level > 7 || nrhs < 1 || !(tol > 0.0) || !out) return CFA_E_ARG;
    GPU_CHECK(cfa_init(0));
    kernel_timings().clear();
    auto t0 = Clock::now();
    const Mesh2D mesh = uniform_mesh(level);
    const double mesh_s = ms_since(t0) / 1000.0;
    const int n = mesh.num_nodes;
    std::vector<double> zero(n, 0.0);
    DeviceMeshBuffers g;
    g.sync_from_host(mesh, zero, zero, zero, zero, zero);
    const cfa_stream_t s = solver_stream();
    GPU_CHECK(cfa_lumped_mass(g.view(), g.d_mass_lumped, s));
    GPU_CHECK(cfa_assemble_pressure(g.view(), g.d_u, g.d_v, g.d_p, g.d_grad_dp_x, g.d_grad_dp_y, g.d_mass_lumped,
                                    g.d_elem_s, Sim::DT_DEFAULT, Sim::RHO, 0.0, g.d_Kp_val, g.d_rhs_p, g.d_div_num,
                                    g.d_dp, nullptr, g.d_inv_m, s));
    AmgPreconditioner amg;
    t0 = Clock::now();
    amg.setup(n, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
    const double setup_cold_ms = host_ms(t0);
    // the steady state of a time-stepping run: buffers already sized, kernels loaded
    double cold_phases[AmgPreconditioner::PH_COUNT];
    for (int k = 0; k < AmgPreconditioner::PH_COUNT; ++k) cold_phases[k] = amg.phase_ms[k];
    t0 = Clock::now();
    amg.setup(n, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
    const double setup_ms = host_ms(t0);
    for (int k = 0; k < AmgPreconditioner::PH_COUNT; ++k) g_amg_phases[k] = amg.phase_ms[k] - cold_phases[k];
    g_amg_phases[AmgPreconditioner::PH_COUNT] = amg.device_levels;
    g_amg_phases[AmgPreconditioner::PH_COUNT + 1] = (double)amg.front_rounds;
    g_amg_phases[AmgPreconditioner::PH_COUNT + 2] = amg.sorted_products;
    GpuLinearSolver solver;
    double solve_ms = 0.0, worst = 0.0;
    long long iters = 0;
    std::vector<double> b(n);
    if (g_keep) {
        keep_pattern(g_kept[0], mesh, g.d_Kp_val, g.nnz);
        g_kept[0].cols = nrhs;
    }
    auto make_rhs = [&](int k) {
        std::mt19937_64 rng(1000 + k);
        std::normal_distribution<double> N(0.0, 1.0);
        double mean = 0.0;
        int free_rows = 0;
        for (int i = 0; i < n; ++i) {
            b[i] = N(rng);
            if (mesh.bc_tag[i] != BC_OUTLET) { mean += b[i]; ++free_rows; }
        }
        mean /= std::max(free_rows, 1);
        for (int i = 0; i < n; ++i) b[i] = mesh.bc_tag[i] == BC_OUTLET ? 0.0 : b[i] - mean;
    };
    for (int k = 0; k < nrhs; ++k) {
        make_rhs(k);
        g.d_rhs_p.upload(b);
        dev_zero(g.d_dp, n * sizeof(double));
        dev_sync();
        t0 = Clock::now();
        const int it = solver.solve_pcg_fused(n, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000, tol,
                                              nullptr, &amg);
        iters += it;
        solve_ms += host_ms(t0);
        worst = std::max(worst, solver.last_rel_residual);
        if (g_keep) {
            KeptSystem& kp = g_kept[0];
            kp.b.insert(kp.b.end(), b.begin(), b.end());
            kp.x0.resize(kp.b.size(), 0.0);
            const std::vector<double> x = g.d_dp.to_host(n);
            kp.x.insert(kp.x.end(), x.begin(), x.end());
            kp.iters.push_back(it);
        }
    }
    // the same right-hand sides with the FP32 cycle (FP32 copies of a second hierarchy; coarse
    // solve, Krylov vectors and operator FP64), as --precision mixed runs the pressure solve
    AmgPreconditioner amg32;
    amg32.fp32 = true;
    amg32.setup(n, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
    double solve32_ms = 0.0, worst32 = 0.0;
    long long iters32 = 0;
    for (int k = 0; k < nrhs; ++k) {
        make_rhs(k);
        g.d_rhs_p.upload(b);
        dev_zero(g.d_dp, n * sizeof(double));
        dev_sync();
        t0 = Clock::now();
        iters32 += solver.solve_pcg_fused(n, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000, tol,
                                          nullptr, &amg32);
        solve32_ms += host_ms(t0);
        worst32 = std::max(worst32, solver.last_rel_residual);
    }
    if (kernel_reps > 0) {
        DevArray<double> t_out;
        t_out.reserve(n);
        const double nnz = g.nnz, nn = n;
        kernel_timings().push_back({"cfa_spmv, 1 right-hand side", 12.0 * nnz + 20.0 * nn + 4.0, time_launches(kernel_reps, [&] {
            GPU_CHECK(cfa_spmv(n, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_dp, t_out, 1, s));
        }), kernel_reps});
        if (!amg.levels.empty()) {
            // one smoothing sweep per level with each lanes-per-row choice of the CSR kernel (the
            // set-up picks one by the mean row length and the level size; these timings are what
            // that rule is checked against), and on the level's sliced-ELL copy when it has one
            cfa_amg_t v = *amg.view();
            std::vector<cfa_amg_level_t> lv(v.levels, v.levels + v.n_levels);
            v.levels = lv.data();
            for (int l = 0; l < v.n_levels; ++l) {
                const cfa_amg_level_t& L = amg.view()->levels[l];
                const double nl = L.rows, nnzl = L.a_nnz;
                const std::string head = "jacobi level " + std::to_string(l) + " rows " + std::to_string(L.rows) +
                                         " nnz/row " + std::to_string((int)std::lround(nnzl / nl));
                lv[l].s_sp = nullptr;
                for (int tpr = 1; tpr <= 32; tpr *= 2) {
                    lv[l].tpr_a = tpr;
                    const std::string name =
                        head + " tpr " + std::to_string(tpr) + (!L.s_sp && tpr == L.tpr_a ? " (chosen)" : "");
                    kernel_timings().push_back({name, 12.0 * nnzl + 36.0 * nl + 4.0, time_launches(kernel_reps, [&] {
                        GPU_CHECK(cfa_amg_smooth(&v, l, L.x, L.b, L.ax, s));
                    }), kernel_reps});
                }
                lv[l].tpr_a = L.tpr_a;
                lv[l].s_sp = L.s_sp;
                if (L.s_sp)
                    kernel_timings().push_back({head + " sliced ELL (chosen)", 12.0 * nnzl + 36.0 * nl + 4.0,
                                                time_launches(kernel_reps, [&] {
                        GPU_CHECK(cfa_amg_smooth(&v, l, L.x, L.b, L.ax, s));
                    }), kernel_reps});
            }
        }
        {   // a device copy of 1 GiB (four times the Infinity Cache): every byte crosses to HBM, so
            // its time is a measured HBM rate and the hardware byte counters of the other kernels
            // are calibrated against it
            const size_t words = (size_t)1 << 27;
            DevArray<double> c_src, c_dst;
            c_src.reserve(words, 0);
            c_dst.reserve(words, 0);
            dev_zero(c_src, words * sizeof(double));
            kernel_timings().push_back({"copy 1 GiB (counter calibration)", 2.0 * words * sizeof(double),
                                        time_launches(kernel_reps, [&] { dev_copy(c_dst, c_src, words); }),
                                        kernel_reps});
        }
        if (!amg32.levels.empty()) {   // the FP32 cycle's level-0 sweep (values, vectors in FP32)
            auto& L = *amg32.levels[0];
            const double n0 = L.A.rows, nnz0 = L.A.nnz;
            dev_zero(L.x32, (size_t)L.A.rows * sizeof(float));
            dev_zero(L.b32, (size_t)L.A.rows * sizeof(float));
            kernel_timings().push_back({L.has_sell ? "level-0 Jacobi sweep, FP32 (k_sell_op, sliced ELL)"
                                                   : "level-0 Jacobi sweep, FP32 (k_csr_op)",
                                        8.0 * nnz0 + 20.0 * n0 + 4.0,
                                        time_launches(kernel_reps, [&] {
                GPU_CHECK(cfa_amg_smooth_f32(amg32.view(), 0, L.x32, L.b32, L.Ax32, s));
            }), kernel_reps});
        }
        // the level-0 sweep as the V-cycle runs it, timed last so that a profiler's last
        // kernel_reps dispatches of this kernel are exactly these
        if (!amg.levels.empty()) {
            const double n0 = amg.levels[0]->A.rows, nnz0 = amg.levels[0]->A.nnz;
            kernel_timings().push_back({amg.levels[0]->has_sell ? "level-0 Jacobi sweep (k_sell_op, sliced ELL)"
                                                                : "level-0 Jacobi sweep (k_csr_op)",
                                        12.0 * nnz0 + 36.0 * n0 + 4.0,
                                        time_launches(kernel_reps, [&] {
                GPU_CHECK(cfa_amg_smooth(amg.view(), 0, g.d_dp, g.d_rhs_p, t_out, s));
            }), kernel_reps});
        }
    }
    const double v[16] = {(double)n, (double)g.nnz, (double)g.n_elems, mesh_s, setup_ms, (double)amg.num_levels(),
                          (double)nrhs, solve_ms, (double)iters, solve_ms / nrhs,
                          iters ? 1000.0 * solve_ms / iters : 0.0, worst, setup_cold_ms,
                          solve32_ms / nrhs, (double)iters32, worst32};
    for (int k = 0; k < 16; ++k) out[k] = v[k];
    return CFA_OK;
}

// Momentum workload: the coupled BDF2 momentum matrix (dt 1e-3, nu 1e-3, SUPG time scale 1e-3,
// skew 0.5, u^n = u^{n-1} = w, p = 0, inflow peak 1.5) for the smooth divergence-free field
// w = curl psi, psi = sum_{k<16} a_k sin(pi m_k x / L) sin(pi l_k y / H) with m_k = 1 + k mod 4,
// l_k = 1 + k / 4 and a_k ~ N(0,1) from mt19937_64(42), scaled to max |w| = 1.5 over the nodes.
// BiCGSTAB2 from x = w to `tol`, once with FP64 and once with FP32 ILU(0) factors.
// out[16] = n, nnz, ne, mesh s, colouring ms, colours, FP64: factor ms, solve ms, it_u, it_v;
// FP32: factor ms, solve ms, it_u, it_v; max |x32 - x64| / max |x64|, us per FP64 iteration.
int cfa_workload_amg_phases(double* out) {
    if (!out) return CFA_E_ARG;
    for (int k = 0; k < AmgPreconditioner::PH_COUNT + 3; ++k) out[k] = g_amg_phases[k];
    return CFA_OK;
}

int cfa_workload_momentum(int level, double tol, int kernel_reps, double* out) {
    if (level < 0 || level > 7 || !(tol > 0.0) || !out) return CFA_E_ARG;
    GPU_CHECK(cfa_init(0));
    kernel_timings().clear();
    auto t0 = Clock::now();
    const Mesh2D mesh = uniform_mesh(level);
    const double mesh_s = ms_since(t0) / 1000.0;
    const int n = mesh.num_nodes;
    std::mt19937_64 rng(42);
    std::normal_distribution<double> N(0.0, 1.0);
    double amp[16];
    for (double& a : amp) a = N(rng);
    std::vector<double> wu(n), wv(n), zero(n, 0.0);
    double wmax = 0.0;
    for (int i = 0; i < n; ++i) {
        double dpx = 0.0, dpy = 0.0;
        for (int k = 0; k < 16; ++k) {
            const double kx = Sim::PI * (1 + k % 4) / Sim::L_CHANNEL, ky = Sim::PI * (1 + k / 4) / Sim::H_CHANNEL;
            dpx += amp[k] * kx * std::cos(kx * mesh.x[i]) * std::sin(ky * mesh.y[i]);
            dpy += amp[k] * ky * std::sin(kx * mesh.x[i]) * std::cos(ky * mesh.y[i]);
        }
        wu[i] = dpy;
        wv[i] = -dpx;
        wmax = std::max(wmax, std::hypot(wu[i], wv[i]));
    }
    for (int i = 0; i < n; ++i) {
        wu[i] *= 1.5 / wmax;
        wv[i] *= 1.5 / wmax;
    }
    DeviceMeshBuffers g;
    g.sync_from_host(mesh, wu, wv, zero, wu, wv);
    const cfa_stream_t s = solver_stream();
    const double dt = Sim::DT_DEFAULT;
    GPU_CHECK(cfa_assemble_momentum(g.view(), g.d_u_old, g.d_v_old, g.d_u_prev, g.d_v_prev, g.d_p, Sim::NU, Sim::RHO,
                                    1.5 / dt, 2.0 / dt, -0.5 / dt, 2.0, -1.0, 1.0e-3, 0.5, nullptr, nullptr, Sim::U_MAX,
                                    -1.0, g.d_A_val, g.d_rhs_u, g.d_rhs_v, g.d_mass_lumped, g.d_u, g.d_v, s));
    DevArray<double> x0u, x0v;
    x0u.reserve(n);
    x0v.reserve(n);
    dev_copy(x0u, g.d_u, n);
    dev_copy(x0v, g.d_v, n);
    if (g_keep) {
        KeptSystem& km = g_kept[1];
        keep_pattern(km, mesh, g.d_A_val, g.nnz);
        km.cols = 2;
        km.b = g.d_rhs_u.to_host(n);
        const std::vector<double> bv = g.d_rhs_v.to_host(n), x0u_h = x0u.to_host(n), x0v_h = x0v.to_host(n);
        km.b.insert(km.b.end(), bv.begin(), bv.end());
        km.x0 = x0u_h;
        km.x0.insert(km.x0.end(), x0v_h.begin(), x0v_h.end());
    }
    MulticolourIlu0 ilu;
    ilu.keep_fp32 = true;
    t0 = Clock::now();
    ilu.set_pattern(n, mesh.csr_row_ptr, mesh.csr_col_ind, g.d_row_ptr, g.d_col_ind);
    const double colour_ms = host_ms(t0);
    double res[2][4];
    std::vector<double> x64, x32;
    for (int prec = 0; prec < 2; ++prec) {
        dev_copy(g.d_u, x0u, n);
        dev_copy(g.d_v, x0v, n);
        t0 = Clock::now();
        ilu.factor(g.d_A_val);
        const double factor_ms = host_ms(t0);
        GpuBicgstab2 sv;
        sv.precond_fp32 = prec == 1;
        t0 = Clock::now();
        sv.solve(n, g.d_row_ptr, g.d_col_ind, g.d_A_val, g.d_rhs_u, g.d_rhs_v, g.d_u, g.d_v, 150, tol, ilu);
        const double solve_ms = host_ms(t0);
        res[prec][0] = factor_ms;
        res[prec][1] = solve_ms;
        res[prec][2] = sv.iters[0];
        res[prec][3] = sv.iters[1];
        std::vector<double>& xs = prec ? x32 : x64;
        xs = g.d_u.to_host(n);
        const std::vector<double> xv = g.d_v.to_host(n);
        xs.insert(xs.end(), xv.begin(), xv.end());
    }
    if (g_keep) {
        g_kept[1].x = x64;
        g_kept[1].x.insert(g_kept[1].x.end(), x32.begin(), x32.end());
        g_kept[1].iters = {(int)res[0][2], (int)res[0][3], (int)res[1][2], (int)res[1][3]};
    }
    double diff = 0.0, scale = 0.0;
    for (size_t i = 0; i < x64.size(); ++i) {
        diff = std::max(diff, std::fabs(x32[i] - x64[i]));
        scale = std::max(scale, std::fabs(x64[i]));
    }
    if (kernel_reps > 0) {
        DevArray<double> t_in, t_out;
        t_in.reserve(2 * (size_t)n);
        t_out.reserve(2 * (size_t)n);
        dev_copy(t_in, g.d_rhs_u, n);
        dev_copy(t_in + n, g.d_rhs_v, n);
        const double nnz = g.nnz, nn = n;
        kernel_timings().push_back({"cfa_spmv, 2 right-hand sides", 12.0 * nnz + 36.0 * nn + 4.0, time_launches(kernel_reps, [&] {
            GPU_CHECK(cfa_spmv(n, g.d_row_ptr, g.d_col_ind, g.d_A_val, t_in, t_out, 2, s));
        }), kernel_reps});
        kernel_timings().push_back({"ilu0 apply, 2 components (all colours)", 12.0 * nnz + 100.0 * nn,
                                    time_launches(kernel_reps, [&] { ilu.apply(t_in, t_out, 2); }), kernel_reps});
    }
    const double its64 = res[0][2] + res[0][3];
    const double v[16] = {(double)n, (double)g.nnz, (double)g.n_elems, mesh_s, colour_ms, (double)ilu.n_colours,
                          res[0][0], res[0][1], res[0][2], res[0][3], res[1][0], res[1][1], res[1][2], res[1][3],
                          scale > 0.0 ? diff / scale : 0.0, its64 > 0 ? 1000.0 * res[0][1] / std::max(res[0][2], res[0][3]) : 0.0};
    for (int k = 0; k < 16; ++k) out[k] = v[k];
    return CFA_OK;
}

// Keep host copies of the systems of the following workload calls (on != 0) or drop them.
int cfa_workload_keep_systems(int on) {
    g_keep = on != 0;
    if (!g_keep) g_kept[0] = g_kept[1] = KeptSystem();
    return CFA_OK;
}

// Sizes of the kept system `which` (0 pressure, 1 momentum): n, nnz, right-hand sides.
int cfa_workload_system_sizes(int which, int* n, int* nnz, int* cols) {
    if (which < 0 || which > 1 || !n || !nnz || !cols) return CFA_E_ARG;
    *n = g_kept[which].n;
    *nnz = g_kept[which].nnz;
    *cols = g_kept[which].cols;
    return CFA_OK;
}

// Copies the kept system: rp[n+1], ci[nnz], tag[n], val[nnz], b and x0 [cols n], x [cols n]
// (pressure: the solution of each right-hand side) or [4 n] (momentum: u, v with FP64 factors,
// then u, v with FP32 factors), and the iterations of each solution in iters.
int cfa_workload_system(int which, int* rp, int* ci, int* tag, double* val, double* b, double* x0, double* x,
                        int* iters) {
    if (which < 0 || which > 1 || !rp || !ci || !tag || !val || !b || !x0 || !x || !iters) return CFA_E_ARG;
    const KeptSystem& k = g_kept[which];
    if (k.n == 0) return CFA_E_NOTREADY;
    std::copy(k.rp.begin(), k.rp.end(), rp);
    std::copy(k.ci.begin(), k.ci.end(), ci);
    std::copy(k.tag.begin(), k.tag.end(), tag);
    std::copy(k.val.begin(), k.val.end(), val);
    std::copy(k.b.begin(), k.b.end(), b);
    std::copy(k.x0.begin(), k.x0.end(), x0);
    std::copy(k.x.begin(), k.x.end(), x);
    std::copy(k.iters.begin(), k.iters.end(), iters);
    return CFA_OK;
}

}  // extern "C"
