// ORACLE (test infrastructure only): the fixed-mesh time loop of src/cylinder_fem_amr.cu:169-538
// (set-up, Steps A-D, MMS errors at the end) on top of the serial kernels and solvers.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "cfa_oracle.hpp"

namespace oracle {

using cfa::Sim;

namespace {
// src/cfa/device.cuh:466-486
void initial_condition(const Mesh2D& m, Vec& u, Vec& v, Vec& p, double peak) {
    u.assign(m.num_nodes, 0.0);
    v.assign(m.num_nodes, 0.0);
    p.assign(m.num_nodes, 0.0);
    for (int i = 0; i < m.num_nodes; ++i) {
        const double yi = m.y[i];
        const double prof = 4.0 * peak * yi * (Sim::H_CHANNEL - yi) / (Sim::H_CHANNEL * Sim::H_CHANNEL);
        if (m.bc_tag[i] == cfa::BC_INLET) {
            u[i] = prof;
        } else if (m.bc_tag[i] == cfa::BC_INTERIOR || m.bc_tag[i] == cfa::BC_OUTLET) {
            const double r = std::hypot(m.x[i] - Sim::X_CYL, yi - Sim::Y_CYL);
            const double damp = 1.0 - std::exp(-std::max(0.0, r - Sim::R_CYL) / 0.03);
            u[i] = damp * prof;
        }
    }
}

double wrap_angle(double a) {
    while (a <= -Sim::PI) a += 2.0 * Sim::PI;
    while (a > Sim::PI) a -= 2.0 * Sim::PI;
    return a;
}

struct Probe { int a = 0, b = 0; double w = 0.0; };
// src/cfa/diagnostics.cuh:23-45
Probe find_probe(const Mesh2D& m, double theta) {
    Probe best;
    double best_dist = 1.0e300;
    for (const auto& ed : m.cylinder_edges) {
        const int a = ed.first, b = ed.second;
        const double ta = std::atan2(m.y[a] - Sim::Y_CYL, m.x[a] - Sim::X_CYL);
        const double tb = std::atan2(m.y[b] - Sim::Y_CYL, m.x[b] - Sim::X_CYL);
        const double span = wrap_angle(tb - ta), s = wrap_angle(theta - ta);
        if (std::abs(span) < 1.0e-14) continue;
        const double w = s / span;
        if (w >= -1.0e-12 && w <= 1.0 + 1.0e-12) return Probe{a, b, std::min(1.0, std::max(0.0, w))};
        const double dist = std::min(std::abs(s), std::abs(wrap_angle(theta - tb)));
        if (dist < best_dist) {
            best_dist = dist;
            best = Probe{a, b, (std::abs(s) < std::abs(wrap_angle(theta - tb))) ? 0.0 : 1.0};
        }
    }
    return best;
}
// Solver-study dump of one step's systems, the layout written by src/cylinder_fem_amr.cu:331-352
// (momentum: int32 n, nnz; row_ptr, col_ind int32; values, rhs_u, rhs_v, u0, v0, x, y float64;
// bc_tag int32; the solutions u, v appended after the solve, :375-378) and :426-438 (pressure:
// int32 n, nnz; row_ptr, col_ind; K_p values, rhs, initial guess).
void dump_raw(std::FILE* f, const void* data, size_t bytes) {
    if (f && bytes) std::fwrite(data, 1, bytes, f);
}
}  // namespace

RunResult run_fixed(const RunConfig& cfg) {
    const auto t_wall0 = cfa::Clock::now();

    // mesh set-up as in src/cylinder_fem_amr.cu:169-207
    Mesh2D mesh;
    cfa::generate_cylinder_mesh(mesh, 8, 7, 28);
    cfa::finalize_mesh(mesh);
    const double inflow0 = cfg.case_2d3 ? 0.0 : Sim::U_MAX;
    Vec hu, hv, hp, huo, hvo;
    initial_condition(mesh, hu, hv, hp, inflow0);
    huo = hu;
    hvo = hv;
    NvbForest amr(mesh, hu, hv, hp);
    amr.extract(mesh, hu, hv, hp, huo, hvo);
    cfa::finalize_mesh(mesh);
    if (cfg.uniform_level > 0 || cfg.min_level > 0) {
        amr.refine_uniform(std::max(cfg.uniform_level, cfg.min_level));
        amr.extract(mesh, hu, hv, hp, huo, hvo);
        cfa::finalize_mesh(mesh);
        initial_condition(mesh, hu, hv, hp, inflow0);
        huo = hu;
        hvo = hv;
    }
    if (cfg.mms)
        for (int i = 0; i < mesh.num_nodes; ++i) {
            const MmsValues r = mms_exact(mesh.x[i], mesh.y[i], 0.0);
            hu[i] = huo[i] = r.u;
            hv[i] = hvo[i] = r.v;
            hp[i] = r.p;
        }

    return run_loop(cfg, mesh, hu, hv, hp, huo, hvo, Vec(mesh.num_nodes, 0.0), 1, t_wall0);
}

RunResult run_state_file(const RunConfig& cfg, const std::string& path) {
    const auto t0 = cfa::Clock::now();
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "[oracle] cannot open %s\n", path.c_str());
        return RunResult();
    }
    int hdr[4] = {0, 0, 0, 0};
    double dt_file = 0.0;
    bool ok = std::fread(hdr, sizeof(int), 4, f) == 4 && hdr[0] == 0x43464153 && hdr[1] > 0 && hdr[2] > 0 &&
              std::fread(&dt_file, sizeof(double), 1, f) == 1;
    const int n = ok ? hdr[1] : 0, ne = ok ? hdr[2] : 0;
    Mesh2D mesh;
    mesh.num_nodes = n;
    mesh.num_elems = ne;
    mesh.x.resize(n);
    mesh.y.resize(n);
    mesh.bc_tag.resize(n);
    std::vector<int> ev(4 * (size_t)ne);
    Vec fields[6];
    ok = ok && std::fread(mesh.x.data(), sizeof(double), n, f) == (size_t)n &&
         std::fread(mesh.y.data(), sizeof(double), n, f) == (size_t)n &&
         std::fread(mesh.bc_tag.data(), sizeof(int), n, f) == (size_t)n &&
         std::fread(ev.data(), sizeof(int), ev.size(), f) == ev.size();
    for (Vec& a : fields) {
        a.resize(n);
        ok = ok && std::fread(a.data(), sizeof(double), n, f) == (size_t)n;
    }
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "[oracle] %s is not a state file\n", path.c_str());
        return RunResult();
    }
    mesh.elems.resize(ne);
    for (int e = 0; e < ne; ++e) {
        for (int k = 0; k < 3; ++k) mesh.elems[e].v[k] = ev[4 * (size_t)e + k];
        mesh.elems[e].gen = ev[4 * (size_t)e + 3];
        mesh.elems[e].level = mesh.elems[e].gen / 2;
    }
    cfa::finalize_mesh(mesh);
    return run_loop(cfg, mesh, fields[0], fields[1], fields[2], fields[3], fields[4], fields[5], hdr[3] + 1, t0);
}

RunResult run_loop(const RunConfig& cfg, Mesh2D& mesh, Vec hu, Vec hv, Vec hp, Vec huo, Vec hvo, Vec dp0,
                   int first_step, cfa::Clock::time_point t_wall0) {
    RunResult res;
    const double dt = cfg.dt;
    const double rot_coeff = cfg.rotational ? Sim::RHO * cfg.nu : 0.0;
    const double force_scale = -2.0 / (Sim::U_MEAN * Sim::U_MEAN * 2.0 * Sim::R_CYL);
    const double surface_scale = -force_scale;
    const int n = mesh.num_nodes, ne = mesh.num_elems, nnz = mesh.nnz;
    std::vector<int> ev(3 * (size_t)ne);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) ev[3 * (size_t)e + k] = mesh.elems[e].v[k];
    const int* rp = mesh.csr_row_ptr.data();
    const int* ci = mesh.csr_col_ind.data();
    const int* tag = mesh.bc_tag.data();
    const double* x = mesh.x.data();
    const double* y = mesh.y.data();
    std::vector<int> cyl(3 * mesh.cylinder_edges.size());
    for (size_t k = 0; k < mesh.cylinder_edges.size(); ++k) {
        cyl[3 * k] = mesh.cylinder_edges[k].first;
        cyl[3 * k + 1] = mesh.cylinder_edges[k].second;
        cyl[3 * k + 2] = mesh.cylinder_edge_elem[k];
    }
    const Probe pf = find_probe(mesh, Sim::PI), pb = find_probe(mesh, 0.0);

    // device upload (src/cfa/device.cuh:131-213): with the outlet parabola, the pinned node, the
    // flux factor and flux-balanced u, u_old; u_prev, v_prev start as copies of u_old, v_old
    Outlet outlet;
    if (cfg.outlet_dirichlet) {
        outlet = outlet_dirichlet_setup(mesh);
        hu = flux_balanced(mesh, hu);
        huo = flux_balanced(mesh, huo);
    }
    Vec u = hu, v = hv, p = hp, u_old = huo, v_old = hvo, u_prev = huo, v_prev = hvo;
    Vec dp = dp0, vort(n, 0.0), A(nnz), Kp(nnz), rhs_u(n), rhs_v(n), rhs_p(n), mass(n);
    Vec gx(n), gy(n), vort_num(n), div_num(n), elem_s(ne), oss_val(4 * (size_t)nnz), inv_m(n);
    Ilu0 ilu;
    ilu.set_pattern(n, rp, ci);
    Amg amg;
    bool amg_dirty = true;
    int amg_built = 0;

    const int last_step = first_step + cfg.steps - 1;
    for (int step = first_step; step <= last_step; ++step) {
        const bool bdf2 = cfg.bdf2 && step > 1;
        const double dt_proj = bdf2 ? (2.0 / 3.0) * dt : dt;
        const double a0 = bdf2 ? 1.5 / dt : 1.0 / dt;
        const double a1 = bdf2 ? 2.0 / dt : 1.0 / dt;
        const double a2 = bdf2 ? -0.5 / dt : 0.0;
        const double c1 = bdf2 ? 2.0 : 1.0;
        const double c2 = bdf2 ? -1.0 : 0.0;
        u_prev = u_old;  v_prev = v_old;
        u_old = u;       v_old = v;

        // Step A: momentum
        std::fill(A.begin(), A.end(), 0.0);
        std::fill(rhs_u.begin(), rhs_u.end(), 0.0);
        std::fill(rhs_v.begin(), rhs_v.end(), 0.0);
        std::fill(mass.begin(), mass.end(), 0.0);
        if (cfg.mms) mms_forcing(n, x, y, step * dt, gx.data(), gy.data());
        assemble_momentum(ne, ev.data(), x, y, u.data(), v.data(), u_prev.data(), v_prev.data(), p.data(),
                          cfg.nu, Sim::RHO, a0, a1, a2, c1, c2,
                          cfg.supg_tau_time > 0.0 ? cfg.supg_tau_time : dt, 0.5, rp, ci, A.data(),
                          rhs_u.data(), rhs_v.data(), mass.data(), cfg.mms ? gx.data() : nullptr,
                          cfg.mms ? gy.data() : nullptr);
        apply_bc_momentum(n, tag, x, y, cfg.u_max > 0.0 ? cfg.u_max : cfa::inflow_peak(step * dt, cfg.case_2d3),
                          cfg.mms ? step * dt : -1.0,
                          rp, ci, A.data(), rhs_u.data(), rhs_v.data(), u.data(), v.data(), outlet);
        const bool dump_now = step == cfg.dump_systems && !cfg.prefix.empty();
        std::FILE* dump = dump_now ? std::fopen((cfg.prefix + "_momentum.bin").c_str(), "wb") : nullptr;
        if (dump) {
            const int hdr[2] = {n, nnz};
            dump_raw(dump, hdr, sizeof(hdr));
            dump_raw(dump, rp, (n + 1) * sizeof(int));
            dump_raw(dump, ci, nnz * sizeof(int));
            dump_raw(dump, A.data(), nnz * sizeof(double));
            for (const double* a : {(const double*)rhs_u.data(), (const double*)rhs_v.data(),
                                    (const double*)u.data(), (const double*)v.data(), x, y})
                dump_raw(dump, a, n * sizeof(double));
            dump_raw(dump, tag, n * sizeof(int));
        }
        ilu.factor(A.data());
        const Bicg2Info bi = bicgstab2(n, rp, ci, A.data(), rhs_u.data(), rhs_v.data(), u.data(), v.data(),
                                       150, cfg.mom_tol, ilu);
        if (dump) {
            dump_raw(dump, u.data(), n * sizeof(double));
            dump_raw(dump, v.data(), n * sizeof(double));
            std::fclose(dump);
        }
        res.nonconverged += (bi.converged[0] ? 0 : 1) + (bi.converged[1] ? 0 : 1);

        // Step B: pressure
        std::fill(Kp.begin(), Kp.end(), 0.0);
        std::fill(rhs_p.begin(), rhs_p.end(), 0.0);
        std::fill(div_num.begin(), div_num.end(), 0.0);
        std::fill(gx.begin(), gx.end(), 0.0);
        std::fill(gy.begin(), gy.end(), 0.0);
        const bool oss = cfg.pressure_stab > 0.0;
        if (oss) {
            compute_oss_weights(ne, ev.data(), x, y, u.data(), v.data(), dt_proj, Sim::RHO, cfg.nu,
                                cfg.pressure_stab, elem_s.data());
            accumulate_scalar_gradient(ne, ev.data(), x, y, p.data(), elem_s.data(), gx.data(), gy.data());
        }
        assemble_pressure_poisson(ne, ev.data(), x, y, u.data(), v.data(), p.data(), gx.data(), gy.data(),
                                  mass.data(), elem_s.data(), dt_proj, Sim::RHO, cfg.pressure_stab, rp, ci,
                                  Kp.data(), rhs_p.data(), div_num.data());
        if (outlet.dirichlet) subtract_mean(n, rhs_p.data(), mass.data());   // fem_kernels.cuh:522-529
        apply_bc_pressure(n, tag, rp, ci, Kp.data(), rhs_p.data(), dp.data(), outlet);
        if (dump_now) {
            if (std::FILE* f = std::fopen((cfg.prefix + "_pressure.bin").c_str(), "wb")) {
                const int hdr[2] = {n, nnz};
                dump_raw(f, hdr, sizeof(hdr));
                dump_raw(f, rp, (n + 1) * sizeof(int));
                dump_raw(f, ci, nnz * sizeof(int));
                dump_raw(f, Kp.data(), nnz * sizeof(double));
                dump_raw(f, rhs_p.data(), n * sizeof(double));
                dump_raw(f, dp.data(), n * sizeof(double));
                std::fclose(f);
            }
        }
        if (oss) {
            std::fill(oss_val.begin(), oss_val.end(), 0.0);
            assemble_oss_operator(ne, ev.data(), x, y, elem_s.data(), rp, ci, nnz, oss_val.data());
            for (int i = 0; i < n; ++i) inv_m[i] = 1.0 / std::fmax(mass[i], 1.0e-14);
        }
        Op oss_op = [&](const double* in, double* out) {
            oss_apply(n, rp, ci, nnz, oss_val.data(), inv_m.data(), tag, in, out, outlet);
        };
        if (amg_dirty || step - amg_built >= cfg.amg_rebuild) {
            Csr K;
            K.rows = K.cols = n;
            K.rp = mesh.csr_row_ptr;
            K.ci = mesh.csr_col_ind;
            K.v = Kp;
            amg.setup(K);
            amg_dirty = false;
            amg_built = step;
        }
        Op prec = [&](const double* r, double* z) { amg.apply(r, z); };
        const SolveInfo pi = pcg(n, rp, ci, Kp.data(), rhs_p.data(), dp.data(), 4000, cfg.ppe_tol,
                                 oss ? &oss_op : nullptr, &prec);
        res.nonconverged += pi.converged ? 0 : 1;
        if (cfg.rotational) {
            std::fill(div_num.begin(), div_num.end(), 0.0);
            stiffness_apply(ne, ev.data(), x, y, dp.data(), -dt_proj / Sim::RHO, div_num.data());
        }

        // Step C: correction
        std::fill(gx.begin(), gx.end(), 0.0);
        std::fill(gy.begin(), gy.end(), 0.0);
        std::fill(vort_num.begin(), vort_num.end(), 0.0);
        accumulate_grad_dp_and_vorticity(ne, ev.data(), x, y, dp.data(), u.data(), v.data(), gx.data(),
                                         gy.data(), vort_num.data());
        velocity_pressure_correction(n, tag, dt_proj, Sim::RHO, 1.0, rot_coeff, div_num.data(), mass.data(),
                                     gx.data(), gy.data(), vort_num.data(), dp.data(), u.data(), v.data(),
                                     p.data(), vort.data(), outlet);

        // Step D: forces and probes
        double diag[6] = {0, 0, 0, 0, 0, 0};
        volume_forces(ne, ev.data(), tag, x, y, u.data(), v.data(), u_old.data(), v_old.data(), u_prev.data(),
                      v_prev.data(), p.data(), a0, -a1, -a2, cfg.nu, Sim::RHO, diag);
        diag[2] = (1.0 - pf.w) * p[pf.a] + pf.w * p[pf.b];
        diag[3] = (1.0 - pb.w) * p[pb.a] + pb.w * p[pb.b];
        surface_forces((int)mesh.cylinder_edges.size(), cyl.data(), ev.data(), x, y, u.data(), v.data(), p.data(),
                       cfg.nu, Sim::RHO, Sim::X_CYL, Sim::Y_CYL, diag + 4);
        StepRow row;
        row.step = step;
        row.t = step * dt;
        row.cd = force_scale * diag[0];
        row.cl = force_scale * diag[1];
        row.dp = diag[2] - diag[3];
        row.nodes = n;
        row.elems = ne;
        row.p_front = diag[2];
        row.p_back = diag[3];
        row.cd_s = surface_scale * diag[4];
        row.cl_s = surface_scale * diag[5];
        row.it_u = bi.iters[0];
        row.it_v = bi.iters[1];
        row.it_p = pi.iters;
        res.rows.push_back(row);
        if (!std::isfinite(row.cd) || !std::isfinite(row.cl) || !std::isfinite(row.dp)) {
            std::fprintf(stderr, "[Step %d] non-finite solution detected, aborting\n", step);
            break;
        }
        if (!cfg.quiet && (step % cfg.print_interval == 0 || step == first_step || step == last_step))
            std::printf("[oracle step %5d] Iters(u,v,p): (%d,%d,%d) | C_D = %.10g | C_L = %.10g | dP = %.10g\n",
                        step, row.it_u, row.it_v, row.it_p, row.cd, row.cl, row.dp);
    }
    res.wall_s = cfa::ms_since(t_wall0) / 1000.0;

    if (cfg.mms) {
        const double t_end = last_step * dt;
        double e2[6] = {0, 0, 0, 0, 0, 0};
        mms_errors(ne, ev.data(), x, y, u.data(), v.data(), p.data(), t_end, e2);
        for (int k = 0; k < 6; ++k) res.mms[k] = std::sqrt(e2[k]);
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == cfa::BC_CYLINDER)
                res.cyl_max = std::max(res.cyl_max, std::fabs(p[i] - mms_exact(x[i], y[i], t_end).p));
        auto probe_err = [&](const Probe& pr, double px, double py) {
            return std::fabs((1.0 - pr.w) * p[pr.a] + pr.w * p[pr.b] - mms_exact(px, py, t_end).p);
        };
        res.probe_max = std::max(probe_err(pf, 0.15, 0.20), probe_err(pb, 0.25, 0.20));
        for (const auto& tri : mesh.elems)
            for (int k = 0; k < 3; ++k) {
                const int a = tri.v[k], b = tri.v[(k + 1) % 3];
                res.h_max = std::max(res.h_max, std::hypot(x[a] - x[b], y[a] - y[b]));
            }
    }

    if (!cfg.prefix.empty()) {
        FILE* f = std::fopen((cfg.prefix + "_forces.csv").c_str(), "w");
        if (f) {
            std::fprintf(f, "step,time,C_D,C_L,delta_p,nodes,elements,p_front,p_back,C_D_surface,C_L_surface\n");
            for (const StepRow& r : res.rows)
                std::fprintf(f, "%d,%.10g,%.10g,%.10g,%.10g,%d,%d,%.10g,%.10g,%.10g,%.10g\n", r.step, r.t, r.cd,
                             r.cl, r.dp, r.nodes, r.elems, r.p_front, r.p_back, r.cd_s, r.cl_s);
            std::fclose(f);
        }
        if (cfg.mms) {
            FILE* g = std::fopen((cfg.prefix + "_mms.csv").c_str(), "w");
            if (g) {
                std::fprintf(g, "nodes,elements,h_max,dt,t,err_u_L2,err_u_H1,err_p_L2,norm_u_L2,norm_u_H1,norm_p_L2,"
                                "err_p_cyl_max,err_p_probe_max\n");
                std::fprintf(g, "%d,%d,%.10g,%.10g,%.10g", n, ne, res.h_max, dt, last_step * dt);
                for (int k = 0; k < 6; ++k) std::fprintf(g, ",%.10g", res.mms[k]);
                std::fprintf(g, ",%.10g,%.10g\n", res.cyl_max, res.probe_max);
                std::fclose(g);
            }
        }
    }
    return res;
}

}  // namespace oracle
