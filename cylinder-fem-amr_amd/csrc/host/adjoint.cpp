// Discrete adjoint of the BDF2 projection step on a frozen mesh, with step checkpointing.
//
// Forward: the step of src/cylinder_fem_amr.cu:277-538 with the reference's default options
// (BDF2 after the first step, rotational incremental projection, SUPG time scale 1e-3, skew 0.5,
// OSS stabilisation 0.3, ILU(0)-BiCGSTAB2 momentum, SA-AMG PCG pressure rebuilt every 50 steps,
// do-nothing outlet), with two parameters theta = (viscosity nu, inflow peak velocity U_max),
// the same device calls in the same order as the program's time loop.
//
// Functional: J = (1/W) sum over the W steps of a window of (w_D C_D + w_L C_L), the
// volume-integral force coefficients of each step (fem_kernels.cuh:1046-1123 scaled by
// -2 / (U_mean^2 D)). The window starts from a given state; the state is
// S^n = (u^n, v^n, u^{n-1}, v^{n-1}, p^n, dp^n), dp being the pressure solver's warm start (no
// influence on the exact step, zero adjoint).
//
// Reverse sweep of one step S^n -> S^{n+1} (u* the momentum solution, dp the pressure increment):
//   forces        : u^{n+1}, u^n, u^{n-1}, p^{n+1}, nu      (cfa_adj_forces)
//   p / u update  : dp_bar, u*_bar = u^{n+1}_bar, p^n_bar += p^{n+1}_bar, nu  (cfa_adj_correction)
//   pressure solve: Op lambda = dp_bar (Op self-adjoint on the free rows, the same PCG + SA-AMG),
//                   then u*, p^n, nu                        (cfa_adj_pressure)
//   momentum solve: A^T lambda = u*_bar (A^T values gathered on the symmetric pattern, its own
//                   ILU(0), BiCGSTAB2), then u^n, u^{n-1}, p^n, nu, U_max  (cfa_adj_momentum)
// Checkpoints every C steps are kept on the device; each segment is recomputed forward, keeping
// its states and momentum solutions, then swept backwards.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "device.hpp"

namespace cfa {
namespace {

constexpr double kTauTime = 1.0e-3;   // SUPG time scale (src/cfa/options.cuh:19)
constexpr double kSkew = 0.5;
constexpr double kStab = 0.3;         // OSS pressure stabilisation (--pressure-stab default)
constexpr int kAmgRebuild = 50;
constexpr int kModes = 64;            // radial modes of the hole boundary

// Radial deformation of the hole boundary r(phi) = R (1 + sum_k eps_k g_k(phi)), g_k Gaussians
// centred at 2 pi k / 64 with width 2 pi / 64 (periodic), carried into the mesh by the smooth
// radial cut-off chi(r) = 1 - t^2 (3 - 2 t), t = (r - R) / (2 R) clamped to [0, 1]: nodes at
// r >= 3 R (the channel walls and the inlet) do not move. The displacement is linear in eps;
// mode k moves node (x, y) by R g_k(phi) chi(r) (x - xc, y - yc) / r.
double mode_weight(int k, double dx, double dy) {
    const double r = std::hypot(dx, dy);
    const double R = Sim::R_CYL, R2 = 3.0 * Sim::R_CYL;
    if (r >= R2 || r <= 0.0) return 0.0;
    const double t = std::min(1.0, std::max(0.0, (r - R) / (R2 - R)));
    const double chi = 1.0 - t * t * (3.0 - 2.0 * t);
    const double phi = std::atan2(dy, dx);
    const double sigma = 2.0 * Sim::PI / kModes;
    double d = phi - 2.0 * Sim::PI * k / kModes;
    d = std::remainder(d, 2.0 * Sim::PI);
    return R * std::exp(-0.5 * (d / sigma) * (d / sigma)) * chi / r;   // times (dx, dy)
}

struct Coeffs {
    double dt_proj, a0, a1, a2, c1, c2;
};
Coeffs coeffs(int step, double dt) {
    const bool bdf2 = step > 1;
    return {bdf2 ? (2.0 / 3.0) * dt : dt, bdf2 ? 1.5 / dt : 1.0 / dt, bdf2 ? 2.0 / dt : 1.0 / dt,
            bdf2 ? -0.5 / dt : 0.0, bdf2 ? 2.0 : 1.0, bdf2 ? -1.0 : 0.0};
}

// CSR position of (j, i) for each (i, j) of a structurally symmetric pattern
std::vector<int> transpose_map(const std::vector<int>& rp, const std::vector<int>& ci) {
    const int n = (int)rp.size() - 1;
    std::vector<int> perm(ci.size(), -1);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            const int* b = ci.data() + rp[j];
            const int* e = ci.data() + rp[j + 1];
            const int* it = std::lower_bound(b, e, i);
            if (it == e || *it != i) return {};
            perm[k] = (int)(it - ci.data());
        }
    return perm;
}

}  // namespace

class AdjointWindow {
public:
    int n = 0, ne = 0, nnz = 0;
    double dt = 1.0e-3, tol = 1.0e-12;
    double force_scale = -2.0 / (Sim::U_MEAN * Sim::U_MEAN * 2.0 * Sim::R_CYL);
    double last_forward_ms = 0.0, last_reverse_ms = 0.0;
    int last_fwd_steps = 0, last_recomputed_steps = 0;
    long long adj_iters_p = 0, adj_iters_m = 0;

    bool init(int level, double dt_, double tol_) {
        dt = dt_;
        tol = tol_;
        generate_cylinder_mesh(mesh_, 8, 7, 28);
        finalize_mesh(mesh_);
        std::vector<double> u, v, p, uo, vo;
        set_initial_condition(mesh_, u, v, p, Sim::U_MAX);
        uo = u;
        vo = v;
        AdaptiveMesh amr(mesh_, u, v, p);
        amr.extract(mesh_, u, v, p, uo, vo);
        finalize_mesh(mesh_);
        if (level > 0) {
            amr.refine_uniform(level);
            amr.extract(mesh_, u, v, p, uo, vo);
            finalize_mesh(mesh_);
        }
        set_initial_condition(mesh_, u, v, p, Sim::U_MAX);
        uo = u;
        vo = v;
        g_.sync_from_host(mesh_, u, v, p, uo, vo);
        x0_ = mesh_.x;
        y0_ = mesh_.y;
        n = g_.n_nodes;
        ne = g_.n_elems;
        nnz = g_.nnz;
        const std::vector<int> perm = transpose_map(mesh_.csr_row_ptr, mesh_.csr_col_ind);
        if ((int)perm.size() != nnz) return false;
        d_perm_.upload(perm);
        ilu_.set_pattern(n, mesh_.csr_row_ptr, mesh_.csr_col_ind, g_.d_row_ptr, g_.d_col_ind);
        d_AT_.reserve(nnz);
        d_ws_.reserve(cfa_adj_workspace_doubles(n, ne));
        d_theta_bar_.reserve(2);
        d_state_.reserve(6 * (size_t)n);
        for (auto* a : {&ustar_, &lam_m_, &lam_p_, &dpb_, &usb_, &vec_tmp_}) a->reserve(2 * (size_t)n);
        for (auto* a : {&adj_a_, &adj_b_}) a->reserve(5 * (size_t)n);
        d_xb_.reserve(n);
        d_yb_.reserve(n);
        initial_[0] = u;
        initial_[1] = v;
        initial_[2] = uo;
        initial_[3] = vo;
        initial_[4] = p;
        return true;
    }

    void initial_state(double* out) const {
        for (int f = 0; f < 5; ++f) std::memcpy(out + (size_t)f * n, initial_[f].data(), n * sizeof(double));
        std::memset(out + 5 * (size_t)n, 0, n * sizeof(double));
    }

    // deform the mesh by the radial modes eps[64] (from the undeformed coordinates)
    void set_shape(const double* eps) {
        std::vector<double> x = x0_, y = y0_;
        for (int k = 0; k < kModes; ++k) eps_[k] = eps[k];
        for (int i = 0; i < n; ++i) {
            const double dx = x0_[i] - Sim::X_CYL, dy = y0_[i] - Sim::Y_CYL;
            double w = 0.0;
            for (int k = 0; k < kModes; ++k)
                if (eps_[k] != 0.0) w += eps_[k] * mode_weight(k, dx, dy);
            x[i] += w * dx;
            y[i] += w * dy;
        }
        g_.d_x.upload(x);
        g_.d_y.upload(y);
        dev_sync();
    }
    // dJ/deps_k = sum_i x_bar_i dx_i/deps_k + y_bar_i dy_i/deps_k
    void shape_gradient(double* out) const {
        const std::vector<double> xb = d_xb_.to_host(n), yb = d_yb_.to_host(n);
        for (int k = 0; k < kModes; ++k) {
            double acc = 0.0;
            for (int i = 0; i < n; ++i) {
                const double dx = x0_[i] - Sim::X_CYL, dy = y0_[i] - Sim::Y_CYL;
                const double w = mode_weight(k, dx, dy);
                if (w != 0.0) acc += w * (xb[i] * dx + yb[i] * dy);
            }
            out[k] = acc;
        }
    }

    // device state [u | v | u_old | v_old | p | dp] <-> the working arrays
    void load(const double* s) {
        double* dst[6] = {g_.d_u, g_.d_v, g_.d_u_old, g_.d_v_old, g_.d_p, g_.d_dp};
        for (int f = 0; f < 6; ++f) dev_copy(dst[f], s + (size_t)f * n, n);
    }
    void store(double* s) const {
        const double* src[6] = {g_.d_u, g_.d_v, g_.d_u_old, g_.d_v_old, g_.d_p, g_.d_dp};
        for (int f = 0; f < 6; ++f) dev_copy(s + (size_t)f * n, src[f], n);
    }
    double* state_buffer() { return d_state_; }

    // One step on the working arrays (the program's Steps A-D). ustar (device, 2n) receives the
    // momentum solution when given. Returns C_D and C_L.
    void step(const double* theta, int step_index, double* ustar, double* cd, double* cl) {
        const double nu = theta[0], umax = theta[1];
        const Coeffs c = coeffs(step_index, dt);
        const cfa_mesh_t* m = g_.view();
        const cfa_stream_t s = solver_stream();
        dev_copy(g_.d_u_prev, g_.d_u_old, n);
        dev_copy(g_.d_v_prev, g_.d_v_old, n);
        dev_copy(g_.d_u_old, g_.d_u, n);
        dev_copy(g_.d_v_old, g_.d_v, n);
        GPU_CHECK(cfa_assemble_momentum(m, g_.d_u_old, g_.d_v_old, g_.d_u_prev, g_.d_v_prev, g_.d_p, nu, Sim::RHO, c.a0,
                                        c.a1, c.a2, c.c1, c.c2, kTauTime, kSkew, nullptr, nullptr, umax, -1.0,
                                        g_.d_A_val, g_.d_rhs_u, g_.d_rhs_v, g_.d_mass_lumped, g_.d_u, g_.d_v, s));
        ilu_.factor(g_.d_A_val);
        mom_.solve(n, g_.d_row_ptr, g_.d_col_ind, g_.d_A_val, g_.d_rhs_u, g_.d_rhs_v, g_.d_u, g_.d_v, 150, tol, ilu_);
        if (ustar) {
            dev_copy(ustar, g_.d_u, n);
            dev_copy(ustar + n, g_.d_v, n);
        }
        GPU_CHECK(cfa_oss_weights(m, g_.d_u, g_.d_v, c.dt_proj, Sim::RHO, nu, kStab, g_.d_elem_s, s));
        GPU_CHECK(cfa_scalar_gradient(m, g_.d_p, g_.d_elem_s, g_.d_grad_dp_x, g_.d_grad_dp_y, s));
        GPU_CHECK(cfa_assemble_pressure(m, g_.d_u, g_.d_v, g_.d_p, g_.d_grad_dp_x, g_.d_grad_dp_y, g_.d_mass_lumped,
                                        g_.d_elem_s, c.dt_proj, Sim::RHO, kStab, g_.d_Kp_val, g_.d_rhs_p,
                                        g_.d_div_num, g_.d_dp, g_.d_oss_val, g_.d_inv_m, s));
        maybe_rebuild_amg(step_index);
        const cfa_oss_op_t op = g_.oss_op();
        solver_.solve_pcg_fused(n, g_.d_row_ptr, g_.d_col_ind, g_.d_Kp_val, g_.d_rhs_p, g_.d_dp, 4000, tol, &op, &amg_);
        GPU_CHECK(cfa_stiffness_apply(m, g_.d_dp, -c.dt_proj / Sim::RHO, g_.d_div_num, s));
        GPU_CHECK(cfa_correction(m, c.dt_proj, Sim::RHO, 1.0, Sim::RHO * nu, g_.d_div_num, g_.d_mass_lumped, g_.d_dp,
                                 g_.d_u, g_.d_v, g_.d_p, g_.d_vort, g_.d_grad_dp_x, g_.d_grad_dp_y, g_.d_vort_num, s));
        const int probe[4] = {g_.probe_front.a, g_.probe_front.b, g_.probe_back.a, g_.probe_back.b};
        GPU_CHECK(cfa_forces(m, g_.d_cyl_edges, g_.n_cyl_edges, g_.d_layer, g_.n_layer, g_.d_u, g_.d_v, g_.d_u_old,
                             g_.d_v_old, g_.d_u_prev, g_.d_v_prev, g_.d_p, c.a0, -c.a1, -c.a2, nu, Sim::RHO, probe,
                             g_.probe_front.w, g_.probe_back.w, g_.d_diag, g_.d_ws, s));
        double diag[6];
        g_.d_diag.download(diag, 6);
        *cd = force_scale * diag[0];
        *cl = force_scale * diag[1];
    }

    // Window forward from a device state; J = (1/W) sum (w_D C_D + w_L C_L). With checkpoints,
    // the state before every C-th step is kept (device, 6n each).
    double forward(const double* theta, int first_step, int steps, const double* w, const double* state0,
                   int every, std::vector<DevArray<double>>* checkpoints, double* history) {
        load(state0);
        double J = 0.0;
        for (int k = 0; k < steps; ++k) {
            if (checkpoints && every > 0 && k % every == 0) {
                DevArray<double>& cp = (*checkpoints)[k / every];
                cp.reserve(6 * (size_t)n);
                store(cp);
            }
            double cd, cl;
            step(theta, first_step + k, nullptr, &cd, &cl);
            if (history) {
                history[2 * k] = cd;
                history[2 * k + 1] = cl;
            }
            J += w[0] * cd + w[1] * cl;
        }
        return J / steps;
    }

    // Gradient of the window functional: theta_bar[2] and, when state_bar is given, dJ/dS^0
    // (device, 6n; the dp slot is zero).
    double gradient(const double* theta, int first_step, int steps, int every, const double* w,
                    const double* state0, double* theta_bar, double* state_bar, double* shape_bar = nullptr) {
        shape_grad_ = shape_bar != nullptr;
        if (every <= 0) every = steps;
        const int n_seg = (steps + every - 1) / every;
        std::vector<DevArray<double>> cps(n_seg);
        auto t0 = Clock::now();
        const double J = forward(theta, first_step, steps, w, state0, every, &cps, nullptr);
        dev_sync();
        last_forward_ms = ms_since(t0);
        last_fwd_steps = steps;
        t0 = Clock::now();
        DevArray<double> seg_states, seg_ustar;
        seg_states.reserve((size_t)(every + 1) * 6 * n);
        seg_ustar.reserve((size_t)every * 2 * n);
        dev_zero(d_theta_bar_, 2 * sizeof(double));
        if (shape_grad_) {
            dev_zero(d_xb_, n * sizeof(double));
            dev_zero(d_yb_, n * sizeof(double));
        }
        dev_zero(adj_a_, 5 * (size_t)n * sizeof(double));   // adjoint of S^W
        dev_zero(lam_m_, 2 * (size_t)n * sizeof(double));
        dev_zero(lam_p_, n * sizeof(double));
        double* adj_next = adj_a_;
        double* adj_cur = adj_b_;
        const double wx = w[0] * force_scale / steps, wy = w[1] * force_scale / steps;
        last_recomputed_steps = 0;
        adj_iters_p = adj_iters_m = 0;
        for (int seg = n_seg - 1; seg >= 0; --seg) {
            const int k0 = seg * every, len = std::min(every, steps - k0);
            load(cps[seg]);
            store(seg_states.ptr);
            for (int k = 0; k < len; ++k) {
                double cd, cl;
                step(theta, first_step + k0 + k, seg_ustar + (size_t)k * 2 * n, &cd, &cl);
                store(seg_states + (size_t)(k + 1) * 6 * n);
            }
            last_recomputed_steps += len;
            for (int k = len - 1; k >= 0; --k) {
                dev_zero(adj_cur, 5 * (size_t)n * sizeof(double));
                adjoint_step(theta, first_step + k0 + k, seg_states + (size_t)k * 6 * n, seg_ustar + (size_t)k * 2 * n,
                             seg_states + (size_t)(k + 1) * 6 * n, wx, wy, adj_next, adj_cur);
                std::swap(adj_next, adj_cur);
            }
        }
        d_theta_bar_.download(theta_bar, 2);
        if (shape_grad_) shape_gradient(shape_bar);
        if (state_bar) {
            dev_copy(state_bar, adj_next, 5 * (size_t)n);
            dev_zero(state_bar + 5 * (size_t)n, n * sizeof(double));
        }
        dev_sync();
        last_reverse_ms = ms_since(t0);
        return J;
    }

private:
    Mesh2D mesh_;
    DeviceMeshBuffers g_;
    MulticolourIlu0 ilu_;
    GpuBicgstab2 mom_;
    GpuLinearSolver solver_;
    AmgPreconditioner amg_;
    int amg_built_ = -1000000;
    DevArray<int> d_perm_;
    DevArray<double> d_AT_, d_ws_, d_theta_bar_, d_state_, ustar_, lam_m_, lam_p_, dpb_, usb_, vec_tmp_;
    DevArray<double> adj_a_, adj_b_, d_xb_, d_yb_;
    std::vector<double> initial_[5];
    std::vector<double> x0_, y0_;       // undeformed node coordinates
    double eps_[kModes] = {};            // radial boundary modes in use
    bool shape_grad_ = false;

    void maybe_rebuild_amg(int step_index) {
        if (std::abs(step_index - amg_built_) >= kAmgRebuild) {
            amg_.setup(n, nnz, g_.d_row_ptr, g_.d_col_ind, g_.d_Kp_val);
            amg_built_ = step_index;
        }
    }

    double norm2(const double* x) { return std::sqrt(solver_.dot(n, x, x)); }

    // adjoint of S^k -> S^{k+1}: reads adj_next = (u, v, u_old, v_old, p)_bar of S^{k+1} (consumed),
    // accumulates (u, v, u_old, v_old, p)_bar of S^k into adj_cur (zeroed by the caller)
    void adjoint_step(const double* theta, int step_index, const double* S0, const double* ustar, const double* S1,
                      double wx, double wy, double* adj_next, double* adj_cur) {
        const double nu = theta[0], umax = theta[1];
        const Coeffs c = coeffs(step_index, dt);
        const cfa_mesh_t* m = g_.view();
        const cfa_stream_t s = solver_stream();
        const size_t N = (size_t)n;
        const double *u0 = S0, *v0 = S0 + N, *uo0 = S0 + 2 * N, *vo0 = S0 + 3 * N, *p0 = S0 + 4 * N;
        const double *u1 = S1, *v1 = S1 + N, *p1 = S1 + 4 * N, *dp = S1 + 5 * N;
        const double *us = ustar, *vs = ustar + N;
        double *au1 = adj_next, *av1 = adj_next + N, *auo1 = adj_next + 2 * N, *avo1 = adj_next + 3 * N,
               *ap1 = adj_next + 4 * N;
        double *au0 = adj_cur, *av0 = adj_cur + N, *auo0 = adj_cur + 2 * N, *avo0 = adj_cur + 3 * N,
               *ap0 = adj_cur + 4 * N;
        // forward operators of this step at its recorded states
        dev_copy(g_.d_u_old, u0, N);
        dev_copy(g_.d_v_old, v0, N);
        dev_copy(g_.d_u_prev, uo0, N);
        dev_copy(g_.d_v_prev, vo0, N);
        dev_copy(g_.d_p, p0, N);
        GPU_CHECK(cfa_assemble_momentum(m, g_.d_u_old, g_.d_v_old, g_.d_u_prev, g_.d_v_prev, g_.d_p, nu, Sim::RHO, c.a0,
                                        c.a1, c.a2, c.c1, c.c2, kTauTime, kSkew, nullptr, nullptr, umax, -1.0,
                                        g_.d_A_val, g_.d_rhs_u, g_.d_rhs_v, g_.d_mass_lumped, g_.d_u, g_.d_v, s));
        GPU_CHECK(cfa_oss_weights(m, us, vs, c.dt_proj, Sim::RHO, nu, kStab, g_.d_elem_s, s));
        GPU_CHECK(cfa_scalar_gradient(m, g_.d_p, g_.d_elem_s, g_.d_grad_dp_x, g_.d_grad_dp_y, s));
        GPU_CHECK(cfa_assemble_pressure(m, us, vs, g_.d_p, g_.d_grad_dp_x, g_.d_grad_dp_y, g_.d_mass_lumped,
                                        g_.d_elem_s, c.dt_proj, Sim::RHO, kStab, g_.d_Kp_val, g_.d_rhs_p,
                                        g_.d_div_num, vec_tmp_, g_.d_oss_val, g_.d_inv_m, s));
        GPU_CHECK(cfa_stiffness_apply(m, dp, -c.dt_proj / Sim::RHO, g_.d_div_num, s));
        maybe_rebuild_amg(step_index);

        // forces of this step
        double* xb = shape_grad_ ? d_xb_.ptr : nullptr;
        double* yb = shape_grad_ ? d_yb_.ptr : nullptr;
        GPU_CHECK(cfa_adj_forces(m, u1, v1, u0, v0, uo0, vo0, p1, c.a0, -c.a1, -c.a2, nu, Sim::RHO, wx, wy, d_ws_,
                                 au1, av1, auo1, avo1, auo0, avo0, ap1, xb, yb, d_theta_bar_, s));
        // rotational pressure update and velocity correction
        GPU_CHECK(cfa_adj_correction(m, c.dt_proj, Sim::RHO, nu, g_.d_div_num, g_.d_mass_lumped, dp, ap1, au1, av1,
                                     d_ws_, dpb_, xb, yb, d_theta_bar_, s));
        GPU_CHECK(cfa_axpy(n, 1.0, ap1, ap0, s));
        double* usb = usb_;
        double* vsb = usb_ + N;
        dev_copy(usb, au1, N);
        dev_copy(vsb, av1, N);
        // pressure: Op lambda = dp_bar, solved on the unit-norm right-hand side
        const double sp = norm2(dpb_);
        if (sp > 0.0) {
            GPU_CHECK(cfa_scale(n, 1.0 / sp, dpb_, s));
            GPU_CHECK(cfa_scale(n, 1.0 / sp, lam_p_, s));
            const cfa_oss_op_t op = g_.oss_op();
            adj_iters_p += solver_.solve_pcg_fused(n, g_.d_row_ptr, g_.d_col_ind, g_.d_Kp_val, dpb_, lam_p_, 4000, tol,
                                                   &op, &amg_);
            GPU_CHECK(cfa_scale(n, sp, lam_p_, s));
            GPU_CHECK(cfa_adj_pressure(m, g_.d_elem_s, us, vs, p0, dp, lam_p_, g_.d_mass_lumped, nu, Sim::RHO,
                                       c.dt_proj, kStab, d_ws_, usb, vsb, ap0, xb, yb, d_theta_bar_, s));
        } else {
            dev_zero(lam_p_, N * sizeof(double));
        }
        // momentum: A^T lambda = u*_bar per component on unit-norm right-hand sides
        GPU_CHECK(cfa_gather_values(nnz, d_perm_, g_.d_A_val, d_AT_, s));
        ilu_.factor(d_AT_);
        const double su = norm2(usb), sv = norm2(vsb);
        double* lu = lam_m_;
        double* lv = lam_m_ + N;
        if (su > 0.0 || sv > 0.0) {
            const double fu = su > 0.0 ? su : 1.0, fv = sv > 0.0 ? sv : 1.0;
            GPU_CHECK(cfa_scale(n, 1.0 / fu, usb, s));
            GPU_CHECK(cfa_scale(n, 1.0 / fu, lu, s));
            GPU_CHECK(cfa_scale(n, 1.0 / fv, vsb, s));
            GPU_CHECK(cfa_scale(n, 1.0 / fv, lv, s));
            mom_.solve(n, g_.d_row_ptr, g_.d_col_ind, d_AT_, usb, vsb, lu, lv, 150, tol, ilu_);
            adj_iters_m += std::max(mom_.iters[0], mom_.iters[1]);
            GPU_CHECK(cfa_scale(n, fu, lu, s));
            GPU_CHECK(cfa_scale(n, fv, lv, s));
            GPU_CHECK(cfa_adj_momentum(m, u0, v0, uo0, vo0, p0, us, vs, lu, lv, nu, Sim::RHO, c.a0, c.a1, c.a2, c.c1,
                                       c.c2, kTauTime, kSkew, Sim::H_CHANNEL, d_ws_, au0, av0, auo0, avo0, ap0, xb, yb,
                                       d_theta_bar_, s));
        } else {
            dev_zero(lam_m_, 2 * N * sizeof(double));
        }
        // u^n appears in S^{k+1} as the previous level
        GPU_CHECK(cfa_axpy(n, 1.0, auo1, au0, s));
        GPU_CHECK(cfa_axpy(n, 1.0, avo1, av0, s));
    }
};

}  // namespace cfa

using namespace cfa;

struct cfa_adjoint {
    AdjointWindow w;
};

extern "C" {

cfa_adjoint* cfa_adjoint_create(int level, double dt, double tol) {
    if (level < 0 || level > 6 || !(dt > 0.0) || !(tol > 0.0)) return nullptr;
    if (cfa_init(0) != CFA_OK) return nullptr;
    auto ctx = std::make_unique<cfa_adjoint>();
    if (!ctx->w.init(level, dt, tol)) return nullptr;
    return ctx.release();
}

void cfa_adjoint_destroy(cfa_adjoint* ctx) { delete ctx; }

int cfa_adjoint_sizes(const cfa_adjoint* ctx, int* n, int* ne, int* nnz) {
    if (!ctx) return CFA_E_ARG;
    if (n) *n = ctx->w.n;
    if (ne) *ne = ctx->w.ne;
    if (nnz) *nnz = ctx->w.nnz;
    return CFA_OK;
}

int cfa_adjoint_initial_state(const cfa_adjoint* ctx, double* state) {
    if (!ctx || !state) return CFA_E_ARG;
    ctx->w.initial_state(state);
    return CFA_OK;
}

int cfa_step_fwd(cfa_adjoint* ctx, const double* theta, int step, const double* state_in, double* state_out,
                 double* forces) {
    if (!ctx || !theta || step < 1 || !state_in || !state_out) return CFA_E_ARG;
    AdjointWindow& w = ctx->w;
    const size_t len = 6 * (size_t)w.n;
    double* d = w.state_buffer();
    GPU_CHECK(cfa_memcpy(d, state_in, len * sizeof(double), 0, solver_stream()));
    w.load(d);
    double cd = 0.0, cl = 0.0;
    w.step(theta, step, nullptr, &cd, &cl);
    w.store(d);
    GPU_CHECK(cfa_memcpy(state_out, d, len * sizeof(double), 1, solver_stream()));
    dev_sync();
    if (forces) {
        forces[0] = cd;
        forces[1] = cl;
    }
    return CFA_OK;
}

int cfa_window_forward(cfa_adjoint* ctx, const double* theta, int first_step, int steps, const double* weights,
                       const double* state0, double* state_out, double* J, double* history) {
    if (!ctx || !theta || first_step < 1 || steps < 1 || !weights || !state0 || !J) return CFA_E_ARG;
    AdjointWindow& w = ctx->w;
    const size_t len = 6 * (size_t)w.n;
    double* d = w.state_buffer();
    GPU_CHECK(cfa_memcpy(d, state0, len * sizeof(double), 0, solver_stream()));
    auto t0 = Clock::now();
    *J = w.forward(theta, first_step, steps, weights, d, 0, nullptr, history);
    dev_sync();
    w.last_forward_ms = ms_since(t0);
    if (state_out) {
        w.store(d);
        GPU_CHECK(cfa_memcpy(state_out, d, len * sizeof(double), 1, solver_stream()));
        dev_sync();
    }
    return CFA_OK;
}

int cfa_adjoint_set_shape(cfa_adjoint* ctx, const double* eps, int n_modes) {
    if (!ctx || !eps || n_modes != kModes) return CFA_E_ARG;
    ctx->w.set_shape(eps);
    return CFA_OK;
}

int cfa_window_adjoint(cfa_adjoint* ctx, const double* theta, int first_step, int steps, int checkpoint_every,
                       const double* weights, const double* state0, double* J, double* dJ_dtheta,
                       double* dJ_dshape, double* dJ_dstate0, double* stats) {
    if (!ctx || !theta || first_step < 1 || steps < 1 || checkpoint_every < 0 || !weights || !state0 || !J ||
        !dJ_dtheta)
        return CFA_E_ARG;
    AdjointWindow& w = ctx->w;
    const size_t len = 6 * (size_t)w.n;
    DevArray<double> d0, dbar;
    d0.reserve(len);
    GPU_CHECK(cfa_memcpy(d0, state0, len * sizeof(double), 0, solver_stream()));
    if (dJ_dstate0) dbar.reserve(len);
    *J = w.gradient(theta, first_step, steps, checkpoint_every, weights, d0, dJ_dtheta,
                    dJ_dstate0 ? dbar.ptr : nullptr, dJ_dshape);
    if (dJ_dstate0) {
        GPU_CHECK(cfa_memcpy(dJ_dstate0, dbar, len * sizeof(double), 1, solver_stream()));
        dev_sync();
    }
    if (stats) {
        stats[0] = w.last_forward_ms;
        stats[1] = w.last_reverse_ms;
        stats[2] = w.last_recomputed_steps;
        stats[3] = (double)w.adj_iters_p;
        stats[4] = (double)w.adj_iters_m;
    }
    return CFA_OK;
}

}  // extern "C"
