// GPU parity of the kernel groups of include/cfa_hip.h against the oracle (the serial CPU
// restatement of the reference's kernels and solvers), on the reference's meshes with seeded
// random fields. Assembly-type kernels must match bit for bit (the gfx950 kernels sum each
// entry's element contributions in the oracle's element order, without contraction); kernels with
// block reductions or other summation orders must match to the stated relative tolerance.
//
//   parity <case> [level]    case: momentum pressure correction forces indicator mms oss ilu
//                                  spmv amg sell pcg bicgstab jacobi mixed topology outlet nvb remap reproject
//                                  all
//
// Prints one line per compared quantity and exits 0 when every comparison passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>

#include "../../cylinder-fem-amr_amd/csrc/host/device.hpp"
#include "../../oracle/cfa_oracle.hpp"

using namespace cfa;
using Vec = std::vector<double>;

namespace {

int failures = 0;

// bitwise equality of two arrays
void same_bits(const char* what, const Vec& gpu, const Vec& ref) {
    size_t bad = 0, first = 0;
    for (size_t i = 0; i < ref.size(); ++i)
        if (std::memcmp(&gpu[i], &ref[i], sizeof(double)) != 0 && !(gpu[i] == 0.0 && ref[i] == 0.0)) {
            if (!bad) first = i;
            ++bad;
        }
    std::printf("  %-34s %9zu values  bit-exact mismatches %zu\n", what, ref.size(), bad);
    if (bad) {
        std::printf("    first at %zu: gpu %.17g oracle %.17g\n", first, gpu[first], ref[first]);
        ++failures;
    }
}

void same_ints(const char* what, const std::vector<int>& gpu, const std::vector<int>& ref) {
    size_t bad = gpu.size() == ref.size() ? 0 : 1;
    for (size_t i = 0; i < std::min(gpu.size(), ref.size()); ++i) bad += gpu[i] != ref[i] ? 1 : 0;
    std::printf("  %-34s %9zu values  mismatches %zu\n", what, ref.size(), bad);
    if (bad) ++failures;
}

// max |gpu - ref| / max |ref| <= tol
void close_rel(const char* what, const Vec& gpu, const Vec& ref, double tol) {
    double d = 0.0, s = 0.0;
    for (size_t i = 0; i < ref.size(); ++i) {
        d = std::max(d, std::fabs(gpu[i] - ref[i]));
        s = std::max(s, std::fabs(ref[i]));
    }
    const double rel = s > 0.0 ? d / s : d;
    const bool ok = rel <= tol && std::isfinite(rel);
    std::printf("  %-34s %9zu values  max rel dev %.3e (tol %.0e)%s\n", what, ref.size(), rel, tol, ok ? "" : "  FAIL");
    if (!ok) ++failures;
}

void expect(bool ok, const char* what) {
    std::printf("  %-34s %s\n", what, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

Mesh2D make_mesh(int level) {
    Mesh2D m;
    generate_cylinder_mesh(m, 8, 7, 28);
    finalize_mesh(m);
    Vec z(m.num_nodes, 0.0), u = z, v = z, p = z, uo = z, vo = z;
    AdaptiveMesh amr(m, z, z, z);
    amr.extract(m, u, v, p, uo, vo);
    finalize_mesh(m);
    if (level > 0) {
        amr.refine_uniform(level);
        amr.extract(m, u, v, p, uo, vo);
        finalize_mesh(m);
    }
    return m;
}

Vec smooth_field(const Mesh2D& m, std::mt19937_64& rng, double scale) {
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double a = U(rng), b = U(rng), c = U(rng), d = U(rng);
    Vec f(m.num_nodes);
    for (int i = 0; i < m.num_nodes; ++i)
        f[i] = scale * (a * std::sin(3.0 * m.x[i] + b) * std::cos(7.0 * m.y[i] + c) + d * m.x[i] * m.y[i] +
                        0.05 * U(rng));
    return f;
}

Vec down(const DevArray<double>& d, size_t n) { return d.to_host(n); }
std::vector<int> down_int(const DevArray<int>& d, size_t n) { return d.to_host(n); }

struct Fixture {
    Mesh2D mesh;
    int n, ne, nnz;
    std::vector<int> ev, cyl;
    const int *rp, *ci, *tag;
    const double *x, *y;
    DeviceMeshBuffers g;
    std::mt19937_64 rng{20261202};
    explicit Fixture(int level) : mesh(make_mesh(level)) {
        n = mesh.num_nodes;
        ne = mesh.num_elems;
        nnz = mesh.nnz;
        ev.resize(3 * (size_t)ne);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) ev[3 * e + k] = mesh.elems[e].v[k];
        for (size_t k = 0; k < mesh.cylinder_edges.size(); ++k) {
            cyl.push_back(mesh.cylinder_edges[k].first);
            cyl.push_back(mesh.cylinder_edges[k].second);
            cyl.push_back(mesh.cylinder_edge_elem[k]);
        }
        rp = mesh.csr_row_ptr.data();
        ci = mesh.csr_col_ind.data();
        tag = mesh.bc_tag.data();
        x = mesh.x.data();
        y = mesh.y.data();
        Vec z(n, 0.0);
        g.sync_from_host(mesh, z, z, z, z, z);
        std::printf("mesh: %d nodes, %d elements, nnz %d\n", n, ne, nnz);
    }
    cfa_stream_t s() const { return solver_stream(); }
};

// ---- cases ---------------------------------------------------------------------------------------
void case_momentum(Fixture& f, bool mms) {
    std::printf("[momentum%s]\n", mms ? " + MMS forcing" : "");
    const Vec un = smooth_field(f.mesh, f.rng, 1.5), vn = smooth_field(f.mesh, f.rng, 0.5);
    const Vec um = smooth_field(f.mesh, f.rng, 1.5), vm = smooth_field(f.mesh, f.rng, 0.5);
    const Vec p = smooth_field(f.mesh, f.rng, 2.0);
    const double dt = 1.0e-3, a0 = 1.5 / dt, a1 = 2.0 / dt, a2 = -0.5 / dt, t = 0.0123;
    f.g.d_u_old.upload(un);
    f.g.d_v_old.upload(vn);
    f.g.d_u_prev.upload(um);
    f.g.d_v_prev.upload(vm);
    f.g.d_p.upload(p);
    f.g.d_u.upload(un);
    f.g.d_v.upload(vn);
    Vec fx(f.n, 0.0), fy(f.n, 0.0);
    if (mms) {
        oracle::mms_forcing(f.n, f.x, f.y, t, fx.data(), fy.data());
        f.g.d_grad_dp_x.upload(fx);   // identical forcing input for both sides
        f.g.d_grad_dp_y.upload(fy);
    }
    GPU_CHECK(cfa_assemble_momentum(f.g.view(), f.g.d_u_old, f.g.d_v_old, f.g.d_u_prev, f.g.d_v_prev, f.g.d_p,
                                    Sim::NU, Sim::RHO, a0, a1, a2, 2.0, -1.0, 1.0e-3, 0.5,
                                    mms ? f.g.d_grad_dp_x.ptr : nullptr, mms ? f.g.d_grad_dp_y.ptr : nullptr, 1.5,
                                    mms ? t : -1.0, f.g.d_A_val, f.g.d_rhs_u, f.g.d_rhs_v, f.g.d_mass_lumped, f.g.d_u,
                                    f.g.d_v, f.s()));
    Vec A(f.nnz, 0.0), ru(f.n, 0.0), rv(f.n, 0.0), mass(f.n, 0.0), u = un, v = vn;
    oracle::assemble_momentum(f.ne, f.ev.data(), f.x, f.y, un.data(), vn.data(), um.data(), vm.data(), p.data(),
                              Sim::NU, Sim::RHO, a0, a1, a2, 2.0, -1.0, 1.0e-3, 0.5, f.rp, f.ci, A.data(), ru.data(),
                              rv.data(), mass.data(), mms ? fx.data() : nullptr, mms ? fy.data() : nullptr);
    oracle::apply_bc_momentum(f.n, f.tag, f.x, f.y, 1.5, mms ? t : -1.0, f.rp, f.ci, A.data(), ru.data(), rv.data(),
                              u.data(), v.data());
    same_bits("A (momentum matrix)", down(f.g.d_A_val, f.nnz), A);
    same_bits("lumped mass", down(f.g.d_mass_lumped, f.n), mass);
    // Dirichlet rows hold boundary data; with MMS they are values of the closed-form solution,
    // computed with the device's sin/cos (within an ulp of the host's), elsewhere bit-exact
    const Vec gru = down(f.g.d_rhs_u, f.n), grv = down(f.g.d_rhs_v, f.n), gu = down(f.g.d_u, f.n),
              gv = down(f.g.d_v, f.n);
    Vec in_g[4], in_o[4], bd_g[4], bd_o[4];
    const Vec* gs[4] = {&gru, &grv, &gu, &gv};
    const Vec* os[4] = {&ru, &rv, &u, &v};
    for (int i = 0; i < f.n; ++i) {
        const bool dir = f.tag[i] == 1 || f.tag[i] == 3 || f.tag[i] == 4;
        for (int k = 0; k < 4; ++k) {
            (dir ? bd_g[k] : in_g[k]).push_back((*gs[k])[i]);
            (dir ? bd_o[k] : in_o[k]).push_back((*os[k])[i]);
        }
    }
    const char* names[4] = {"rhs_u", "rhs_v", "u", "v"};
    for (int k = 0; k < 4; ++k) {
        same_bits((std::string(names[k]) + " (free rows)").c_str(), in_g[k], in_o[k]);
        if (mms)
            close_rel((std::string(names[k]) + " (Dirichlet rows, MMS data)").c_str(), bd_g[k], bd_o[k], 1e-15);
        else
            same_bits((std::string(names[k]) + " (Dirichlet rows)").c_str(), bd_g[k], bd_o[k]);
    }
}

// pressure system with OSS; leaves Kp, rhs, oss operator on the device (and returns the oracle's)
struct PressureSystem { Vec Kp, rhs, oss, inv_m, elem_s, mass; };
PressureSystem case_pressure(Fixture& f, bool report = true) {
    if (report) std::printf("[pressure]\n");
    const Vec us = smooth_field(f.mesh, f.rng, 1.5), vs = smooth_field(f.mesh, f.rng, 0.5);
    const Vec p = smooth_field(f.mesh, f.rng, 2.0);
    const double dtp = (2.0 / 3.0) * 1.0e-3, stab = 0.3;
    f.g.d_u.upload(us);
    f.g.d_v.upload(vs);
    f.g.d_p.upload(p);
    const cfa_mesh_t* m = f.g.view();
    GPU_CHECK(cfa_lumped_mass(m, f.g.d_mass_lumped, f.s()));
    GPU_CHECK(cfa_oss_weights(m, f.g.d_u, f.g.d_v, dtp, Sim::RHO, Sim::NU, stab, f.g.d_elem_s, f.s()));
    GPU_CHECK(cfa_scalar_gradient(m, f.g.d_p, f.g.d_elem_s, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.s()));
    const Vec gpx = down(f.g.d_grad_dp_x, f.n), gpy = down(f.g.d_grad_dp_y, f.n);
    GPU_CHECK(cfa_assemble_pressure(m, f.g.d_u, f.g.d_v, f.g.d_p, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.g.d_mass_lumped,
                                    f.g.d_elem_s, dtp, Sim::RHO, stab, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_div_num,
                                    f.g.d_dp, f.g.d_oss_val, f.g.d_inv_m, f.s()));
    PressureSystem o;
    o.mass.assign(f.n, 0.0);
    oracle::assemble_lumped_mass(f.ne, f.ev.data(), f.x, f.y, o.mass.data());
    o.elem_s.assign(f.ne, 0.0);
    oracle::compute_oss_weights(f.ne, f.ev.data(), f.x, f.y, us.data(), vs.data(), dtp, Sim::RHO, Sim::NU, stab,
                                o.elem_s.data());
    Vec gx(f.n, 0.0), gy(f.n, 0.0), div(f.n, 0.0), dp(f.n, 1.0);
    oracle::accumulate_scalar_gradient(f.ne, f.ev.data(), f.x, f.y, p.data(), o.elem_s.data(), gx.data(), gy.data());
    o.Kp.assign(f.nnz, 0.0);
    o.rhs.assign(f.n, 0.0);
    oracle::assemble_pressure_poisson(f.ne, f.ev.data(), f.x, f.y, us.data(), vs.data(), p.data(), gx.data(),
                                      gy.data(), o.mass.data(), o.elem_s.data(), dtp, Sim::RHO, stab, f.rp, f.ci,
                                      o.Kp.data(), o.rhs.data(), div.data());
    oracle::apply_bc_pressure(f.n, f.tag, f.rp, f.ci, o.Kp.data(), o.rhs.data(), dp.data());
    o.oss.assign(4 * (size_t)f.nnz, 0.0);
    oracle::assemble_oss_operator(f.ne, f.ev.data(), f.x, f.y, o.elem_s.data(), f.rp, f.ci, f.nnz, o.oss.data());
    o.inv_m.resize(f.n);
    for (int i = 0; i < f.n; ++i) o.inv_m[i] = 1.0 / std::fmax(o.mass[i], 1.0e-14);
    if (report) {
        same_bits("lumped mass", down(f.g.d_mass_lumped, f.n), o.mass);
        same_bits("OSS weights s_K", down(f.g.d_elem_s, f.ne), o.elem_s);
        same_bits("G(p) x", gpx, gx);
        same_bits("G(p) y", gpy, gy);
        same_bits("Kp (PPE + BCs)", down(f.g.d_Kp_val, f.nnz), o.Kp);
        same_bits("rhs_p", down(f.g.d_rhs_p, f.n), o.rhs);
        same_bits("div_num", down(f.g.d_div_num, f.n), div);
        same_bits("OSS operator [Bx|By|BTx|BTy]", down(f.g.d_oss_val, 4 * (size_t)f.nnz), o.oss);
        same_bits("inverse lumped mass", down(f.g.d_inv_m, f.n), o.inv_m);
        Vec dpg = down(f.g.d_dp, f.n);
        int bad = 0;
        for (int i = 0; i < f.n; ++i) bad += (f.tag[i] == 2 && dpg[i] != 0.0) ? 1 : 0;
        expect(bad == 0, "dp = 0 on outlet rows");
    }
    return o;
}

void case_correction(Fixture& f) {
    std::printf("[correction, rotational term]\n");
    const Vec dp = smooth_field(f.mesh, f.rng, 0.1), u = smooth_field(f.mesh, f.rng, 1.5),
              v = smooth_field(f.mesh, f.rng, 0.5), p = smooth_field(f.mesh, f.rng, 2.0);
    const double dtp = (2.0 / 3.0) * 1.0e-3, rot = Sim::RHO * Sim::NU;
    Vec mass(f.n, 0.0);
    oracle::assemble_lumped_mass(f.ne, f.ev.data(), f.x, f.y, mass.data());
    f.g.d_mass_lumped.upload(mass);
    f.g.d_dp.upload(dp);
    f.g.d_u.upload(u);
    f.g.d_v.upload(v);
    f.g.d_p.upload(p);
    const cfa_mesh_t* m = f.g.view();
    GPU_CHECK(cfa_stiffness_apply(m, f.g.d_dp, -dtp / Sim::RHO, f.g.d_div_num, f.s()));
    Vec div(f.n, 0.0);
    oracle::stiffness_apply(f.ne, f.ev.data(), f.x, f.y, dp.data(), -dtp / Sim::RHO, div.data());
    same_bits("div_num = -(dt/rho) K dp", down(f.g.d_div_num, f.n), div);
    GPU_CHECK(cfa_correction(m, dtp, Sim::RHO, 1.0, rot, f.g.d_div_num, f.g.d_mass_lumped, f.g.d_dp, f.g.d_u, f.g.d_v,
                             f.g.d_p, f.g.d_vort, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.g.d_vort_num, f.s()));
    Vec gx(f.n, 0.0), gy(f.n, 0.0), vn(f.n, 0.0), uu = u, vv = v, pp = p, vort(f.n, 0.0);
    oracle::accumulate_grad_dp_and_vorticity(f.ne, f.ev.data(), f.x, f.y, dp.data(), u.data(), v.data(), gx.data(),
                                             gy.data(), vn.data());
    oracle::velocity_pressure_correction(f.n, f.tag, dtp, Sim::RHO, 1.0, rot, div.data(), mass.data(), gx.data(),
                                         gy.data(), vn.data(), dp.data(), uu.data(), vv.data(), pp.data(),
                                         vort.data());
    same_bits("u corrected", down(f.g.d_u, f.n), uu);
    same_bits("v corrected", down(f.g.d_v, f.n), vv);
    same_bits("p updated", down(f.g.d_p, f.n), pp);
    same_bits("vorticity", down(f.g.d_vort, f.n), vort);
}

void case_forces(Fixture& f) {
    std::printf("[forces and probes]\n");
    Vec fields[7];
    for (int k = 0; k < 7; ++k) fields[k] = smooth_field(f.mesh, f.rng, k < 6 ? 1.0 : 3.0);
    f.g.d_u.upload(fields[0]);
    f.g.d_v.upload(fields[1]);
    f.g.d_u_old.upload(fields[2]);
    f.g.d_v_old.upload(fields[3]);
    f.g.d_u_prev.upload(fields[4]);
    f.g.d_v_prev.upload(fields[5]);
    f.g.d_p.upload(fields[6]);
    const double a0 = 1500.0, a1 = 2000.0, a2 = -500.0;
    const int probe[4] = {f.g.probe_front.a, f.g.probe_front.b, f.g.probe_back.a, f.g.probe_back.b};
    GPU_CHECK(cfa_forces(f.g.view(), f.g.d_cyl_edges, f.g.n_cyl_edges, f.g.d_layer, f.g.n_layer, f.g.d_u, f.g.d_v,
                         f.g.d_u_old, f.g.d_v_old, f.g.d_u_prev, f.g.d_v_prev, f.g.d_p, a0, -a1, -a2, Sim::NU,
                         Sim::RHO, probe, f.g.probe_front.w, f.g.probe_back.w, f.g.d_diag, f.g.d_ws, f.s()));
    const Vec d = down(f.g.d_diag, 6);
    double o[6] = {0, 0, 0, 0, 0, 0};
    oracle::volume_forces(f.ne, f.ev.data(), f.tag, f.x, f.y, fields[0].data(), fields[1].data(), fields[2].data(),
                          fields[3].data(), fields[4].data(), fields[5].data(), fields[6].data(), a0, -a1, -a2,
                          Sim::NU, Sim::RHO, o);
    const Vec& p = fields[6];
    o[2] = (1.0 - f.g.probe_front.w) * p[probe[0]] + f.g.probe_front.w * p[probe[1]];
    o[3] = (1.0 - f.g.probe_back.w) * p[probe[2]] + f.g.probe_back.w * p[probe[3]];
    oracle::surface_forces((int)f.mesh.cylinder_edges.size(), f.cyl.data(), f.ev.data(), f.x, f.y, fields[0].data(),
                           fields[1].data(), p.data(), Sim::NU, Sim::RHO, Sim::X_CYL, Sim::Y_CYL, o + 4);
    close_rel("volume forces (F_x, F_y)", Vec(d.begin(), d.begin() + 2), Vec(o, o + 2), 1e-12);
    same_bits("probe pressures", Vec(d.begin() + 2, d.begin() + 4), Vec(o + 2, o + 4));
    close_rel("surface forces", Vec(d.begin() + 4, d.end()), Vec(o + 4, o + 6), 1e-12);
}

void case_indicator(Fixture& f) {
    std::printf("[AMR error indicator]\n");
    const Vec u = smooth_field(f.mesh, f.rng, 1.5), v = smooth_field(f.mesh, f.rng, 0.5);
    f.g.d_u.upload(u);
    f.g.d_v.upload(v);
    GPU_CHECK(cfa_error_indicator(f.g.view(), f.g.d_u, f.g.d_v, 1.5, f.g.d_mass_lumped, f.g.d_grad_rec,
                                  f.g.d_elem_error, f.s()));
    Vec mass(f.n, 0.0), gr(4 * (size_t)f.n, 0.0), ee(2 * (size_t)f.ne, 0.0);
    oracle::error_indicator(f.ne, f.n, f.ev.data(), f.x, f.y, u.data(), v.data(), 1.5, mass.data(), gr.data(),
                            ee.data());
    same_bits("recovered gradient", down(f.g.d_grad_rec, 4 * (size_t)f.n), gr);
    same_bits("eta_v | eta_i", down(f.g.d_elem_error, 2 * (size_t)f.ne), ee);
}

void case_mms(Fixture& f) {
    std::printf("[MMS forcing and errors]\n");
    const double t = 0.0375;
    GPU_CHECK(cfa_mms_forcing(f.n, f.g.d_x, f.g.d_y, t, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.s()));
    Vec fx(f.n), fy(f.n);
    oracle::mms_forcing(f.n, f.x, f.y, t, fx.data(), fy.data());
    close_rel("MMS forcing f_x (device libm)", down(f.g.d_grad_dp_x, f.n), fx, 1e-14);
    close_rel("MMS forcing f_y (device libm)", down(f.g.d_grad_dp_y, f.n), fy, 1e-14);
    const Vec u = smooth_field(f.mesh, f.rng, 1.0), v = smooth_field(f.mesh, f.rng, 1.0),
              p = smooth_field(f.mesh, f.rng, 1.0);
    f.g.d_u.upload(u);
    f.g.d_v.upload(v);
    f.g.d_p.upload(p);
    DevArray<double> out;
    out.reserve(6, 0);
    GPU_CHECK(cfa_mms_errors(f.g.view(), f.g.d_u, f.g.d_v, f.g.d_p, t, out, f.g.d_ws, f.s()));
    Vec o(6, 0.0);
    oracle::mms_errors(f.ne, f.ev.data(), f.x, f.y, u.data(), v.data(), p.data(), t, o.data());
    close_rel("MMS error integrals", down(out, 6), o, 1e-12);
}

void case_oss(Fixture& f) {
    std::printf("[OSS operator apply]\n");
    const PressureSystem o = case_pressure(f, false);
    const Vec in = smooth_field(f.mesh, f.rng, 1.0);
    DevArray<double> d_in, d_out;
    d_in.upload(in);
    d_out.upload(Vec(f.n, 0.0));
    GPU_CHECK(cfa_oss_apply(f.g.view(), f.g.d_oss_val, f.g.d_inv_m, d_in, d_out, f.g.d_oss_gx, f.g.d_oss_gy, f.s()));
    Vec out(f.n, 0.0);
    oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, o.oss.data(), o.inv_m.data(), f.tag, in.data(), out.data());
    close_rel("-B^T M^-1 B in (assembled)", down(d_out, f.n), out, 1e-13);
    d_out.upload(Vec(f.n, 0.0));
    GPU_CHECK(cfa_oss_apply_matrix_free(f.g.view(), f.g.d_elem_s, f.g.d_mass_lumped, d_in, d_out, f.g.d_oss_gx,
                                        f.g.d_oss_gy, f.s()));
    Vec mf(f.n, 0.0);
    oracle::oss_apply_matrix_free(f.ne, f.n, f.ev.data(), f.x, f.y, f.tag, o.elem_s.data(), o.mass.data(), in.data(),
                                  mf.data());
    close_rel("-B^T M^-1 B in (matrix-free)", down(d_out, f.n), mf, 1e-13);
}

// random non-symmetric diagonally dominant matrix on the mesh pattern (the reference's test)
Vec convection_like(const Fixture& f, std::mt19937_64& rng) {
    std::uniform_real_distribution<double> U(0.0, 1.0);
    Vec a(f.nnz);
    for (int i = 0; i < f.n; ++i) {
        double off = 0.0;
        int d = -1;
        for (int k = f.rp[i]; k < f.rp[i + 1]; ++k) {
            if (f.ci[k] == i) { d = k; continue; }
            a[k] = -(0.2 + U(rng)) * (f.ci[k] > i ? 1.6 : 0.6);
            off += std::fabs(a[k]);
        }
        a[d] = 1.05 * off + 0.1;
    }
    return a;
}

void case_ilu(Fixture& f) {
    std::printf("[multicolour ILU(0)]\n");
    const Vec a = convection_like(f, f.rng);
    f.g.d_A_val.upload(a);
    MulticolourIlu0 ilu;
    ilu.keep_fp32 = true;
    ilu.sell_min_rows = 0;   // the sliced-ELL copies at this size too (by default from 2^20 rows)
    ilu.set_pattern(f.n, f.mesh.csr_row_ptr, f.mesh.csr_col_ind, f.g.d_row_ptr, f.g.d_col_ind);
    ilu.factor(f.g.d_A_val);
    oracle::Ilu0 o;
    o.set_pattern(f.n, f.rp, f.ci);
    o.factor(a.data());
    same_ints("colours (greedy, node order)", ilu.d_colour.to_host(f.n), o.colour);
    expect(ilu.n_colours == o.n_colours, "number of colours");
    same_bits("LU factors (in place, CSR order)", down(ilu.d_lu, f.nnz), o.lu);
    Vec r(2 * (size_t)f.n);
    for (auto& v : r) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
    DevArray<double> d_r, d_z;
    d_r.upload(r);
    d_z.reserve(2 * (size_t)f.n);
    ilu.apply(d_r, d_z, 2);
    Vec z(2 * (size_t)f.n);
    o.apply(r.data(), z.data(), 2);
    close_rel("(LU)^-1 r, two components", down(d_z, 2 * (size_t)f.n), z, 1e-13);
    // the one-lane sweeps on the sliced-ELL rows against the same sweeps on the CSR rows: bit for
    // bit (the rows sum in their CSR order; padding subtracts exact zeros)
    expect(ilu.has_sell, "split rows carry the sliced-ELL copies");
    if (ilu.has_sell) {
        std::printf("  sliced ELL: L %d slots (%d entries), U %d slots (%d entries), longest rows %d / %d\n",
                    ilu.s_lslots, ilu.nl, ilu.s_uslots, ilu.nu, ilu.s_llen, ilu.s_ulen);
        cfa_ilu_t csr = *ilu.view();
        csr.s_lrow = csr.s_urow = nullptr;
        DevArray<double> d_z2;
        d_z2.reserve(2 * (size_t)f.n);
        GPU_CHECK(cfa_ilu0_apply(&csr, d_r, d_z2, 2, f.s()));
        same_bits("(LU)^-1 r: sliced ELL = CSR rows", down(d_z, 2 * (size_t)f.n), down(d_z2, 2 * (size_t)f.n));
        GPU_CHECK(cfa_ilu0_apply(ilu.view(), d_r, d_z, 1, f.s()));
        GPU_CHECK(cfa_ilu0_apply(&csr, d_r, d_z2, 1, f.s()));
        same_bits("(LU)^-1 r, one component: sliced ELL = CSR", down(d_z, f.n), down(d_z2, f.n));
    }
}

void case_spmv(Fixture& f) {
    std::printf("[SpMV and dot]\n");
    const Vec a = convection_like(f, f.rng);
    Vec xv(2 * (size_t)f.n);
    for (auto& v : xv) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
    DevArray<double> d_a, d_x, d_y, ws;
    d_a.upload(a);
    d_x.upload(xv);
    d_y.reserve(2 * (size_t)f.n);
    ws.reserve(2048, 0);
    GPU_CHECK(cfa_spmv(f.n, f.g.d_row_ptr, f.g.d_col_ind, d_a, d_x, d_y, 2, f.s()));
    Vec y(2 * (size_t)f.n);
    oracle::spmv(f.n, f.rp, f.ci, a.data(), xv.data(), y.data());
    oracle::spmv(f.n, f.rp, f.ci, a.data(), xv.data() + f.n, y.data() + f.n);
    close_rel("A x, two right-hand sides", down(d_y, 2 * (size_t)f.n), y, 1e-14);
    double dg = 0.0;
    GPU_CHECK(cfa_dot(f.n, d_x, d_x + f.n, &dg, ws, f.s()));
    const double dor = oracle::dot(f.n, xv.data(), xv.data() + f.n);
    close_rel("dot", Vec{dg}, Vec{dor}, 1e-12);
}

oracle::Csr host_csr(const Fixture& f, const Vec& v) {
    oracle::Csr K;
    K.rows = K.cols = f.n;
    K.rp = f.mesh.csr_row_ptr;
    K.ci = f.mesh.csr_col_ind;
    K.v = v;
    return K;
}

void case_amg(Fixture& f) {
    std::printf("[SA-AMG set-up and V-cycle]\n");
    const PressureSystem o = case_pressure(f, false);
    oracle::Amg oa;
    oa.setup(host_csr(f, o.Kp));
    // the default products, then every product through the global-memory sort that takes rows
    // beyond the LDS kernel
    for (int sorted = 0; sorted < 2; ++sorted) {
        AmgPreconditioner amg;
        amg.keep_aggregates = true;
        amg.force_sorted_products = sorted == 1;
        if (sorted) std::printf("  every product by the sorted global-memory method:\n");
        amg.setup(f.n, f.nnz, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val);
        std::printf("  aggregation decision-front rounds: %lld, sorted products %d\n", amg.front_rounds,
                    amg.sorted_products);
        if (sorted) expect(amg.sorted_products == 3 * (int)amg.levels.size(), "every product sorted");
        expect(amg.num_levels() == oa.num_levels(), "number of levels");
        std::vector<int> rows_g, rows_o, nnz_g, nnz_o;
        Vec om_g, om_o;
        for (size_t l = 0; l < amg.levels.size(); ++l) {
            rows_g.push_back(amg.levels[l]->A.rows);
            nnz_g.push_back(amg.levels[l]->A.nnz);
            nnz_g.push_back(amg.levels[l]->P.nnz);
            om_g.push_back(amg.levels[l]->omega);
        }
        for (const auto& L : oa.levels) {
            rows_o.push_back(L.A.rows);
            nnz_o.push_back(L.A.rp.back());
            nnz_o.push_back(L.P.rp.back());
            om_o.push_back(L.omega);
        }
        same_ints("level rows", rows_g, rows_o);
        same_ints("level nnz (A, P)", nnz_g, nnz_o);
        same_bits("Jacobi weights omega", om_g, om_o);
        expect(amg.coarse_n == oa.coarse_n, "coarse size");
        std::printf("  levels built on the device: %d of %zu\n", amg.device_levels, amg.levels.size());
        // every level's operator, prolongator, restriction and inverse diagonal, bit for bit
        if (amg.levels.size() == oa.levels.size()) {
            std::vector<int> pat_g, pat_o;
            Vec val_g, val_o;
            auto add = [&](const AmgPreconditioner::DevCsr& D, const oracle::Csr& C) {
                const std::vector<int> rp = down_int(D.rp, D.rows + 1), ci = down_int(D.ci, D.nnz);
                pat_g.insert(pat_g.end(), rp.begin(), rp.end());
                pat_g.insert(pat_g.end(), ci.begin(), ci.end());
                pat_o.insert(pat_o.end(), C.rp.begin(), C.rp.end());
                pat_o.insert(pat_o.end(), C.ci.begin(), C.ci.end());
                const Vec v = down(D.v, D.nnz);
                val_g.insert(val_g.end(), v.begin(), v.end());
                val_o.insert(val_o.end(), C.v.begin(), C.v.end());
            };
            for (size_t l = 0; l < amg.levels.size(); ++l) {
                const auto& L = *amg.levels[l];
                const auto& O = oa.levels[l];
                add(L.A, O.A);
                add(L.P, O.P);
                add(L.R, O.R);
                const Vec d = down(L.dinv, L.A.rows);
                val_g.insert(val_g.end(), d.begin(), d.end());
                val_o.insert(val_o.end(), O.dinv.begin(), O.dinv.end());
            }
            same_ints("hierarchy patterns (A, P, R per level)", pat_g, pat_o);
            same_bits("hierarchy values (A, P, R, D^-1 per level)", val_g, val_o);
        }
        if (!oa.levels.empty()) {
            std::vector<int> agg;
            oracle::smoothed_prolongator(oa.levels[0].A, oa.levels[0].dinv, oa.levels[0].omega, 0.08, &agg);
            same_ints("level-0 aggregates", amg.last_aggregates, agg);
        }
        if (amg.coarse_n == oa.coarse_n)
            close_rel("coarse-level inverse", down(amg.d_coarse_inv, (size_t)amg.coarse_n * amg.coarse_n),
                      oa.coarse_inv, 1e-12);
        Vec r(f.n);
        for (auto& v : r) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
        DevArray<double> d_r, d_z;
        d_r.upload(r);
        d_z.reserve(f.n);
        Vec z(f.n);
        oa.apply(r.data(), z.data());
        amg.apply(d_r, d_z);
        close_rel("one V-cycle", down(d_z, f.n), z, 1e-12);
        if (!sorted) {
            // two V-cycles at the same time on two streams, one through the second set of work
            // vectors (the concurrent re-projection solves): each equals the V-cycle run alone
            const Vec z1 = down(d_z, f.n);
            Vec r2(f.n);
            for (auto& v : r2) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
            DevArray<double> d_r2, d_z2, d_za, d_zb;
            d_r2.upload(r2);
            d_z2.reserve(f.n);
            d_za.reserve(f.n);
            d_zb.reserve(f.n);
            amg.apply(d_r2, d_z2);
            const Vec z2 = down(d_z2, f.n);
            const cfa_amg_t* v2 = amg.second_view();
            cfa_stream_t s2 = nullptr;
            GPU_CHECK(cfa_stream_create(&s2));
            for (int rep = 0; rep < 3; ++rep) {
                GPU_CHECK(cfa_amg_apply(v2, d_r2, d_zb, s2));
                GPU_CHECK(cfa_amg_apply(amg.view(), d_r, d_za, solver_stream()));
            }
            GPU_CHECK(cfa_stream_sync(s2));
            dev_sync();
            GPU_CHECK(cfa_stream_destroy(s2));
            same_bits("V-cycle beside a second one", down(d_za, f.n), z1);
            same_bits("second V-cycle (second vectors)", down(d_zb, f.n), z2);
        }
    }
}

// The sliced-ELL copy of the level-0 operator: its layout against the CSR rows (every position's
// row, the window order by decreasing length, every entry and the zero padding); its sweeps against
// the CSR kernel with the level's lanes per row, whose summation order they keep, FP64 and FP32 bit
// for bit; and a whole V-cycle (FP64 and FP32) on it against one on the CSR rows, bit for bit.
void case_sell(Fixture& f) {
    std::printf("[sliced-ELL level-0 operator]\n");
    (void)case_pressure(f, false);
    for (int f32 = 0; f32 < 2; ++f32) {
        AmgPreconditioner amg;
        amg.fp32 = f32 == 1;
        amg.setup(f.n, f.nnz, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val);
        if (amg.levels.empty()) {
            expect(false, "a hierarchy with a level 0");
            return;
        }
        const auto& L = *amg.levels[0];
        // the copy keeps the summation order of 1, 2 or 4 lanes per row; below 20k rows the lanes
        // rule gives more (a lane per entry) and the level keeps the CSR kernels
        const bool want = L.A.tpr <= 4;
        expect(L.has_sell == want, want ? "level 0 carries the sliced-ELL copy"
                                        : "no sliced-ELL copy (more than 4 CSR lanes per row)");
        if (!L.has_sell) return;
        const int n = L.A.rows, positions = (n + 63) / 64 * 64, slices = positions / 64;
        if (!f32) {
            const std::vector<int> rp = down_int(L.A.rp, n + 1), ci = down_int(L.A.ci, L.A.nnz);
            const Vec av = down(L.A.v, L.A.nnz);
            const std::vector<int> srow = down_int(L.s_row, positions), sp = down_int(L.s_sp, slices + 1);
            const std::vector<int> sci = down_int(L.s_ci, std::max(L.s_entries, 1));
            const Vec sv = down(L.s_v, std::max(L.s_entries, 1));
            int longest = 0, bad_perm = 0, bad_order = 0, bad_len = 0, bad_pad = 0;
            std::vector<int> seen(n, 0);
            std::vector<int> ci_g, ci_o;
            Vec v_g, v_o;
            for (int r = 0; r < n; ++r) longest = std::max(longest, rp[r + 1] - rp[r]);
            for (int p = 0; p < positions; ++p) {
                const int r = srow[p], w0 = p / 256 * 256;
                if (r < 0) {   // padding only past the last row
                    bad_perm += p < n ? 1 : 0;
                    continue;
                }
                if (r >= n || r < w0 || r >= w0 + 256 || seen[r]++) {
                    ++bad_perm;
                    continue;
                }
                const int len = rp[r + 1] - rp[r];
                if (p > w0 && srow[p - 1] >= 0) {   // decreasing length, ties by row
                    const int q = srow[p - 1], lq = rp[q + 1] - rp[q];
                    bad_order += (lq < len || (lq == len && q > r)) ? 1 : 0;
                }
                const int sl = (sp[p / 64 + 1] - sp[p / 64]) / 64;
                bad_len += (p % 64 == 0 && sl != len) || sl < len ? 1 : 0;
                for (int k = 0; k < sl; ++k) {
                    const int at = sp[p / 64] + 64 * k + p % 64;
                    if (k < len) {
                        ci_g.push_back(sci[at]);
                        ci_o.push_back(ci[rp[r] + k]);
                        v_g.push_back(sv[at]);
                        v_o.push_back(av[rp[r] + k]);
                    } else {
                        bad_pad += (sci[at] != 0 || sv[at] != 0.0) ? 1 : 0;
                    }
                }
            }
            for (int r = 0; r < n; ++r) bad_perm += seen[r] == 1 ? 0 : 1;
            std::printf("  %d rows, %d slices, %d slots for %d entries, longest row %d (recorded %d)\n", n, slices,
                        sp[slices], L.A.nnz, longest, L.s_len);
            expect(bad_perm == 0, "every row at one position of its window");
            expect(bad_order == 0, "window order: decreasing length, ties by row");
            expect(bad_len == 0 && longest == L.s_len && sp[slices] == L.s_entries, "slice lengths and sizes");
            expect(bad_pad == 0, "padding slots (column 0, value 0)");
            same_ints("column indices in CSR order", ci_g, ci_o);
            same_bits("values in CSR order", v_g, v_o);
        }
        // one sweep on the copy against the CSR kernel with one lane per row
        const cfa_amg_t* view = amg.view();
        cfa_amg_t csr_view = *view;
        std::vector<cfa_amg_level_t> lv(view->levels, view->levels + view->n_levels);
        lv[0].s_sp = nullptr;
        csr_view.levels = lv.data();
        std::printf("  lanes per row of the level-0 CSR kernels: %d\n", lv[0].tpr_a);
        Vec x(n), b(n);
        for (auto& v : x) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
        for (auto& v : b) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
        if (!f32) {
            DevArray<double> d_x, d_b, d_y1, d_y2;
            d_x.upload(x);
            d_b.upload(b);
            d_y1.reserve(n);
            d_y2.reserve(n);
            GPU_CHECK(cfa_amg_smooth(view, 0, d_x, d_b, d_y1, f.s()));
            GPU_CHECK(cfa_amg_smooth(&csr_view, 0, d_x, d_b, d_y2, f.s()));
            same_bits("FP64 sweep: sliced ELL = CSR", down(d_y1, n), down(d_y2, n));
        } else {
            std::vector<float> x32(x.begin(), x.end()), b32(b.begin(), b.end());
            DevArray<float> d_x, d_b, d_y1, d_y2;
            d_x.upload(x32);
            d_b.upload(b32);
            d_y1.reserve(n);
            d_y2.reserve(n);
            GPU_CHECK(cfa_amg_smooth_f32(view, 0, d_x, d_b, d_y1, f.s()));
            GPU_CHECK(cfa_amg_smooth_f32(&csr_view, 0, d_x, d_b, d_y2, f.s()));
            const std::vector<float> y1 = d_y1.to_host(n), y2 = d_y2.to_host(n);
            same_bits("FP32 sweep: sliced ELL = CSR", Vec(y1.begin(), y1.end()), Vec(y2.begin(), y2.end()));
        }
        DevArray<double> d_r, d_z1, d_z2;
        d_r.upload(x);
        d_z1.reserve(n);
        d_z2.reserve(n);
        GPU_CHECK(cfa_amg_apply(view, d_r, d_z1, f.s()));
        GPU_CHECK(cfa_amg_apply(&csr_view, d_r, d_z2, f.s()));
        same_bits(f32 ? "FP32 V-cycle: sliced ELL = CSR" : "FP64 V-cycle: sliced ELL = CSR", down(d_z1, n),
                  down(d_z2, n));
    }
}

// The scheme-consistent re-projection after a remesh with its two solves (u^{n-1} and u^n) side by
// side on two streams against the same two solves one after the other, from the same state: the
// velocities, the vorticity and the last increment bit for bit.
void case_reproject(Fixture& f) {
    std::printf("[re-projection: the two solves side by side]\n");
    const Vec u = smooth_field(f.mesh, f.rng, 1.5), v = smooth_field(f.mesh, f.rng, 0.5);
    const Vec uo = smooth_field(f.mesh, f.rng, 1.5), vo = smooth_field(f.mesh, f.rng, 0.5);
    const Vec p = smooth_field(f.mesh, f.rng, 2.0);
    const double dt_proj = (2.0 / 3.0) * 5.0e-4, stab = 0.3;
    Vec out[2][6];
    int iters[2] = {0, 0};
    for (int run = 0; run < 2; ++run) {
        f.g.d_u.upload(u);
        f.g.d_v.upload(v);
        f.g.d_u_old.upload(uo);
        f.g.d_v_old.upload(vo);
        f.g.d_p.upload(p);
        AmgPreconditioner amg;
        GpuLinearSolver solver;
        ReprojectionAux aux;
        DevArray<double> phi;
        phi.reserve(f.n);
        int nonconverged = 0;
        iters[run] = project_scheme_consistent(f.g, solver, dt_proj, stab, &amg, nonconverged, nullptr, true, phi,
                                               1.0e-8, run ? &aux : nullptr);
        expect(nonconverged == 0, run ? "side by side: both solves converged" : "one after the other: converged");
        out[run][0] = down(f.g.d_u, f.n);
        out[run][1] = down(f.g.d_v, f.n);
        out[run][2] = down(f.g.d_u_old, f.n);
        out[run][3] = down(f.g.d_v_old, f.n);
        out[run][4] = down(f.g.d_vort, f.n);
        out[run][5] = down(phi, f.n);
    }
    std::printf("  PCG iterations (both solves): one after the other %d, side by side %d\n", iters[0], iters[1]);
    expect(iters[0] == iters[1], "same iterations");
    const char* names[6] = {"u^n", "v^n", "u^{n-1}", "v^{n-1}", "vorticity", "last increment"};
    for (int k = 0; k < 6; ++k) same_bits(names[k], out[1][k], out[0][k]);
}

// The no-allocation mean removal against cfa_pressure_remove_mean: on the fixture's mesh, then
// on a smaller one (the second call must not see the first's sums in the shared array), bit for
// bit; a mesh beyond the fixed array is refused before any device work.
void case_noalloc(Fixture& f, int level) {
    std::printf("[pressure mean removal without allocation]\n");
    Mesh2D small = make_mesh(level >= 2 ? level - 2 : 0);
    DeviceMeshBuffers gs;
    gs.outlet_dirichlet = true;
    const Vec zs(small.num_nodes, 0.0), z(f.n, 0.0);
    gs.sync_from_host(small, zs, zs, zs, zs, zs);
    f.g.outlet_dirichlet = true;
    f.g.sync_from_host(f.mesh, z, z, z, z, z);
    struct Case {
        const char* name;
        const cfa_mesh_t* view;
        int n;
    };
    const Case cases[2] = {{"fixture mesh", f.g.view(), f.n}, {"smaller mesh", gs.view(), small.num_nodes}};
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    DevArray<double> d_mass, d_a, d_b, d_ws;
    for (const Case& c : cases) {
        Vec mass(c.n), rhs(c.n);
        for (int i = 0; i < c.n; ++i) {
            mass[i] = 1.0e-4 * (1.5 + U(f.rng));
            rhs[i] = U(f.rng);
        }
        d_mass.upload(mass);
        d_a.upload(rhs);
        d_b.upload(rhs);
        d_ws.reserve(cfa_pressure_mean_ws_doubles(c.n));
        GPU_CHECK(cfa_pressure_remove_mean(c.view, d_mass, d_a, d_ws, f.s()));
        GPU_CHECK(cfa_pressure_remove_mean_noalloc(c.view, d_mass, d_b, f.s()));
        std::printf("  %s: %d nodes\n", c.name, c.n);
        same_bits("  mean-free, pinned rhs (no allocation vs workspace)", down(d_b, c.n), down(d_a, c.n));
    }
    cfa_mesh_t big = *f.g.view();
    big.n_nodes = CFA_MEAN_NOALLOC_MAX_NODES + 1;
    const int rc = cfa_pressure_remove_mean_noalloc(&big, d_mass, d_b, f.s());
    std::printf("  %d nodes: %s\n", big.n_nodes, cfa_error_string(rc));
    expect(rc == CFA_E_SIZE, "a mesh beyond the fixed array is refused");
    f.g.outlet_dirichlet = false;   // back to the default outlet for later cases
    f.g.sync_from_host(f.mesh, z, z, z, z, z);
}

// Mesh topology on the device against the host build: the CSR graph, the incidence and position
// tables, the cylinder edges and the checks, on the fixture's uniform mesh and on a locally
// refined NVB mesh; then on two planted defects (a flipped element, a missing element), which
// both paths must reject.
void topology_pair(const char* name, const Mesh2D& m, bool expect_valid) {
    std::printf("  %s: %d nodes, %d elements\n", name, (int)m.x.size(), (int)m.elems.size());
    Mesh2D h = m;
    build_csr_and_boundary_topology(h);
    MeshQuality qh;
    const bool vh = validate_mesh(h, qh);
    Mesh2D d = m;
    d.csr_row_ptr.clear();
    d.csr_col_ind.clear();
    d.nnz = 0;
    d.cylinder_edges.clear();
    d.cylinder_edge_elem.clear();
    DeviceMeshBuffers g;
    const Vec z(m.x.size(), 0.0);
    MeshQuality qd;
    bool vd = false;
    const bool took = g.sync_from_host_device_topology(d, z, z, z, z, z, qd, vd);
    expect(took, "the device path takes the mesh");
    expect(vh == expect_valid, "host check as planted");
    expect(vd == vh, "device check = host check");
    same_ints("CSR row pointers", d.csr_row_ptr, h.csr_row_ptr);
    same_ints("CSR columns", d.csr_col_ind, h.csr_col_ind);
    std::vector<int> ce_d, ce_h;
    for (size_t k = 0; k < d.cylinder_edges.size(); ++k)
        ce_d.insert(ce_d.end(), {d.cylinder_edges[k].first, d.cylinder_edges[k].second, d.cylinder_edge_elem[k]});
    for (size_t k = 0; k < h.cylinder_edges.size(); ++k)
        ce_h.insert(ce_h.end(), {h.cylinder_edges[k].first, h.cylinder_edges[k].second, h.cylinder_edge_elem[k]});
    same_ints("cylinder edges", ce_d, ce_h);
    same_bits("minimum area, minimum angle", Vec{qd.min_area, qd.min_angle_deg}, Vec{qh.min_area, qh.min_angle_deg});
    // summed in another order: bounded by (elements) x eps = 4.5e-11 at 2e5 elements
    close_rel("total area", Vec{qd.total_area}, Vec{qh.total_area}, 1e-10);
    expect(qd.max_level == qh.max_level && qd.max_gen == qh.max_gen, "maximum level and generation");
    std::vector<int> ip, ie, ipos;
    build_incidence(h, ip, ie, ipos);
    same_ints("incidence pointers", down_int(g.d_inc_ptr, (int)ip.size()), ip);
    same_ints("incidence elements", down_int(g.d_inc_elem, (int)ie.size()), ie);
    same_ints("owner-computes positions", down_int(g.d_inc_pos, (int)ipos.size()), ipos);
    // the tables alone, from the host's graph (the upload of an overlapped remesh)
    DeviceMeshBuffers g2;
    g2.sync_from_host_device_tables(h, z, z, z, z, z);
    same_ints("tables from the host graph: incidence", down_int(g2.d_inc_elem, (int)ie.size()), ie);
    same_ints("tables from the host graph: positions", down_int(g2.d_inc_pos, (int)ipos.size()), ipos);
}

// The device NVB forest against the oracle's restatement of the reference's rules: uniform
// refinement, then adaptation cycles on synthetic indicators (a moving blob, a thin stripe, noise)
// with look-ahead along a travelling-wave velocity, tight element budgets every fourth cycle, one or
// two buffer layers, and the overlapped-remesh transfer of a later state every other cycle.
// Connectivity, generations, boundary tags, coordinates (the cylinder midpoints' arc positions are
// placed by the host's libm), vertex slots and transferred values must match bit for bit.
void compare_meshes(const char* what, const Mesh2D& a, const Mesh2D& b, const Vec* fa, const Vec* fb) {
    std::vector<int> ea, eb;
    for (const Triangle& t : a.elems) ea.insert(ea.end(), {t.v[0], t.v[1], t.v[2], t.gen});
    for (const Triangle& t : b.elems) eb.insert(eb.end(), {t.v[0], t.v[1], t.v[2], t.gen});
    char name[96];
    std::snprintf(name, sizeof name, "%s: elements", what);
    same_ints(name, ea, eb);
    std::snprintf(name, sizeof name, "%s: boundary tags", what);
    same_ints(name, a.bc_tag, b.bc_tag);
    if (a.num_nodes != b.num_nodes) return;
    std::snprintf(name, sizeof name, "%s: coordinates", what);
    Vec ca = a.x, cb = b.x;
    ca.insert(ca.end(), a.y.begin(), a.y.end());
    cb.insert(cb.end(), b.y.begin(), b.y.end());
    same_bits(name, ca, cb);
    static const char* fields[5] = {"u", "v", "p", "u_old", "v_old"};
    for (int q = 0; q < 5; ++q) {
        std::snprintf(name, sizeof name, "%s: %s", what, fields[q]);
        same_bits(name, fa[q], fb[q]);
    }
}

void case_nvb(Fixture&, int level) {
    std::printf("[NVB forest on the device vs the oracle]\n");
    Mesh2D base;
    generate_cylinder_mesh(base, 8, 7, 28);
    finalize_mesh(base);
    const Vec z(base.num_nodes, 0.0);
    AdaptiveMesh P(base, z, z, z);
    oracle::NvbForest O(base, z, z, z);
    Mesh2D mp, mo;
    Vec fp[5], fo[5];
    P.extract(mp, fp[0], fp[1], fp[2], fp[3], fp[4]);
    O.extract(mo, fo[0], fo[1], fo[2], fo[3], fo[4]);
    compare_meshes("base mesh", mp, mo, fp, fo);
    P.refine_uniform(2);
    O.refine_uniform(2);
    P.extract(mp, fp[0], fp[1], fp[2], fp[3], fp[4]);
    O.extract(mo, fo[0], fo[1], fo[2], fo[3], fo[4]);
    compare_meshes("uniform level 2", mp, mo, fp, fo);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int cycles = 4 * level;
    for (int c = 0; c < cycles; ++c) {
        const int n = mp.num_nodes, L = mp.num_elems;
        Vec st[5];
        for (Vec& v : st) v.resize(n);
        for (int i = 0; i < n; ++i) {
            st[0][i] = 1.0 + 0.3 * std::sin(8.0 * mp.x[i] - 0.3 * c) + 0.01 * U(rng);
            st[1][i] = 0.2 * std::cos(5.0 * mp.y[i] + 0.2 * c);
            st[2][i] = U(rng);
            st[3][i] = 0.99 * st[0][i];
            st[4][i] = 0.99 * st[1][i];
        }
        P.scatter_solution(st[0], st[1], st[2], st[3], st[4]);
        O.scatter_solution(st[0], st[1], st[2], st[3], st[4]);
        Vec eta(2 * (size_t)L);
        const double bx = 0.3 + 0.05 * c, by = 0.2 + 0.05 * std::sin(0.7 * c);
        for (int e = 0; e < L; ++e) {
            const int* v = mp.elems[e].v;
            const double xc = (mp.x[v[0]] + mp.x[v[1]] + mp.x[v[2]]) / 3.0;
            const double yc = (mp.y[v[0]] + mp.y[v[1]] + mp.y[v[2]]) / 3.0;
            const double r2 = (xc - bx) * (xc - bx) + (yc - by) * (yc - by);
            eta[e] = 0.4 * std::exp(-r2 / 0.002) + 0.01 * U(rng) +
                     (std::fabs(yc - 0.2 - 0.02 * std::sin(10.0 * xc)) < 0.01 ? 0.08 : 0.0);
            eta[L + e] = 0.03 * std::exp(-r2 / 0.01) + 0.001 * U(rng);
        }
        AmrParams pp;
        oracle::AmrParams po;
        pp.tol_vort = po.tol_vort = 0.05;
        pp.tol_interp = po.tol_interp = 0.005;
        pp.max_level = po.max_level = 5;
        pp.min_level = po.min_level = 2;
        pp.max_elems = po.max_elems = (c % 4 == 3) ? (int)(L * 1.05) : 160000;
        pp.lookahead_time = po.lookahead_time = (c % 3 == 0) ? 0.0 : 0.02 * (1 + c % 5);
        pp.buffer_layers = po.buffer_layers = 1 + (c % 2);
        const AmrStats a = P.adapt(eta, pp);
        const oracle::AmrCounts b = O.adapt(eta, po);
        std::printf("  cycle %2d: marked %d/%d, coarsen %d/%d, bisections %d/%d + closure %d/%d, removed %d/%d, "
                    "leaves %d/%d (device/oracle)\n", c, a.marked_refine, b.marked_refine, a.marked_coarsen,
                    b.marked_coarsen, a.bisections, b.bisections, a.closure_bisections, b.closure_bisections,
                    a.removed_vertices, b.removed_vertices, a.leaves_after, b.leaves_after);
        expect(a.marked_refine == b.marked_refine && a.marked_coarsen == b.marked_coarsen &&
               a.bisections == b.bisections && a.closure_bisections == b.closure_bisections &&
               a.removed_vertices == b.removed_vertices && a.leaves_after == b.leaves_after,
               "adaptation counts");
        {
            const std::vector<int>& a = P.pass_sizes();
            const std::vector<int>& b = O.pass_sizes();
            std::string sa, sb;
            for (int k : a) sa += " " + std::to_string(k);
            for (int k : b) sb += " " + std::to_string(k);
            if (a != b) std::printf("  passes (work > 0, closure < 0): device%s\n                                  oracle%s\n", sa.c_str(), sb.c_str());
            same_ints("bisections per refinement pass", a, b);
        }
        const std::vector<int> old_p = P.node_vertices(), old_o = O.node_slots();
        P.extract(mp, fp[0], fp[1], fp[2], fp[3], fp[4]);
        O.extract(mo, fo[0], fo[1], fo[2], fo[3], fo[4]);
        char what[48];
        std::snprintf(what, sizeof what, "cycle %d", c);
        compare_meshes(what, mp, mo, fp, fo);
        same_ints("element slots of the leaves", P.leaf_elements(), O.leaf_slots());
        same_ints("vertex slots of the nodes", P.node_vertices(), O.node_slots());
        int cnt[4];
        P.slot_counts(cnt);
        std::printf("  slots device {elements %d, vertices %d, free %d, %d} oracle {%d, %d, free %d, %d}\n", cnt[0],
                    cnt[1], cnt[2], cnt[3], O.element_slots(), O.vertex_slots(), O.free_element_slots(),
                    O.free_vertex_slots());
        expect(cnt[0] == O.element_slots() && cnt[1] == O.vertex_slots() && cnt[2] == O.free_element_slots() &&
               cnt[3] == O.free_vertex_slots(), "slot counts");
        if (c % 2) {   // the overlapped remesh's transfer of a later state
            Vec t[5], rp[5], ro[5];
            for (Vec& v : t) v.resize(old_p.size());
            for (size_t i = 0; i < old_p.size(); ++i)
                for (int q = 0; q < 5; ++q) t[q][i] = U(rng);
            P.inlet_umax = O.inlet_umax = 1.2;
            P.retransfer(old_p, t[0], t[1], t[2], t[3], t[4], rp[0], rp[1], rp[2], rp[3], rp[4]);
            O.retransfer(old_o, t[0], t[1], t[2], t[3], t[4], ro[0], ro[1], ro[2], ro[3], ro[4]);
            for (int q = 0; q < 5; ++q) same_bits("transferred later state", rp[q], ro[q]);
            P.inlet_umax = O.inlet_umax = Sim::U_MAX;
        }
        if (failures) break;
    }
}

// AMG remap (src/cfa/linalg.cuh:1074-1157, --amg-remap): the level-0 prolongator of a hierarchy
// built on one mesh carried to the locally refined mesh that follows it, against the oracle's
// restatement applied to the oracle's own prolongator of the same matrix. Matrix: the graph
// Laplacian with a 1e-4 diagonal shift on each mesh. The pattern must match; values to 1e-14 (a
// new node's row sums equal columns in the order its neighbours list them, the reference in the
// order of an unstable sort). Then PCG with the remapped hierarchy converges on the new mesh.
std::vector<double> laplacian_values(const Mesh2D& m) {
    std::vector<double> a(m.nnz);
    for (int i = 0; i < m.num_nodes; ++i) {
        int d = -1, deg = 0;
        for (int k = m.csr_row_ptr[i]; k < m.csr_row_ptr[i + 1]; ++k) {
            if (m.csr_col_ind[k] == i) { d = k; continue; }
            a[k] = -1.0;
            ++deg;
        }
        a[d] = deg * (1.0 + 1.0e-4);
    }
    return a;
}

void case_remap(Fixture& f) {
    std::printf("[AMG remap to a refined mesh]\n");
    Mesh2D m = f.mesh;
    Vec z(m.num_nodes, 0.0), u = z, v = z, p = z, uo = z, vo = z;
    AdaptiveMesh amr(m, z, z, z);
    amr.extract(m, u, v, p, uo, vo);
    finalize_mesh(m);
    const Vec a0 = laplacian_values(m);
    DevArray<int> d_rp, d_ci;
    DevArray<double> d_a;
    d_rp.upload(m.csr_row_ptr);
    d_ci.upload(m.csr_col_ind);
    d_a.upload(a0);
    AmgPreconditioner amg;
    amg.keep_p0_host = true;
    amg.setup(m.num_nodes, m.nnz, d_rp, d_ci, d_a);
    oracle::Csr A0;
    A0.rows = A0.cols = m.num_nodes;
    A0.rp = m.csr_row_ptr;
    A0.ci = m.csr_col_ind;
    A0.v = a0;
    oracle::Amg oa;
    oa.setup(A0);
    const std::vector<int> old_vids = amr.node_vertices();
    const Vec old_x = m.x, old_y = m.y;
    // refine around (0.4, 0.2) and one level everywhere near the outlet
    AmrParams prm;
    prm.max_level = 5;
    prm.min_level = 0;
    prm.keep_radius = 0.0;
    prm.tol_vort = prm.tol_interp = 1.0;
    Vec eta(2 * m.elems.size(), 0.5);
    for (size_t e = 0; e < m.elems.size(); ++e) {
        const Triangle& t = m.elems[e];
        const double cx = (m.x[t.v[0]] + m.x[t.v[1]] + m.x[t.v[2]]) / 3.0;
        const double cy = (m.y[t.v[0]] + m.y[t.v[1]] + m.y[t.v[2]]) / 3.0;
        if (std::hypot(cx - 0.4, cy - 0.2) < 0.03 || cx > 2.0) eta[e] = 8.0;
    }
    amr.adapt(eta, prm);
    amr.extract(m, u, v, p, uo, vo);
    finalize_mesh(m);
    const int n = m.num_nodes;
    std::vector<int> slot_old(amr.num_vertex_slots(), -1), new_to_old(n);
    for (int j = 0; j < (int)old_vids.size(); ++j) slot_old[old_vids[j]] = j;
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        const int j = slot_old[amr.node_vertices()[i]];
        new_to_old[i] = (j >= 0 && old_x[j] == m.x[i] && old_y[j] == m.y[i]) ? j : -1;
        kept += new_to_old[i] >= 0;
    }
    std::printf("  %d -> %d nodes, %d kept\n", (int)old_vids.size(), n, kept);
    const Vec a1 = laplacian_values(m);
    d_rp.upload(m.csr_row_ptr);
    d_ci.upload(m.csr_col_ind);
    d_a.upload(a1);
    amg.remap(n, m.nnz, d_rp, d_ci, d_a, m.csr_row_ptr.data(), m.csr_col_ind.data(), new_to_old);
    expect(amg.remaps == 1, "remapped (no fall-back to a new set-up)");
    bool ok = false;
    const oracle::Csr Pn = oracle::amg_remap_p0(oa.levels[0].P, n, m.csr_row_ptr.data(), m.csr_col_ind.data(),
                                                new_to_old, &ok);
    expect(ok, "oracle remap reached every new node");
    {   // the oracle's remap by what its rule fixes: a kept node keeps its old row exactly; a new
        // node's row (a mean of its neighbours' rows) is nonempty, its columns come from those
        // rows and its sum lies within theirs
        const oracle::Csr& P0 = oa.levels[0].P;
        int kept_bad = 0, new_bad = 0;
        auto row_sum = [&](int i) {
            double t = 0.0;
            for (int a = Pn.rp[i]; a < Pn.rp[i + 1]; ++a) t += Pn.v[a];
            return t;
        };
        for (int i = 0; i < n; ++i) {
            const int j = new_to_old[i];
            if (j >= 0) {
                const bool same = Pn.rp[i + 1] - Pn.rp[i] == P0.rp[j + 1] - P0.rp[j] &&
                                  std::equal(Pn.ci.begin() + Pn.rp[i], Pn.ci.begin() + Pn.rp[i + 1],
                                             P0.ci.begin() + P0.rp[j]) &&
                                  std::equal(Pn.v.begin() + Pn.rp[i], Pn.v.begin() + Pn.rp[i + 1],
                                             P0.v.begin() + P0.rp[j]);
                kept_bad += same ? 0 : 1;
                continue;
            }
            std::vector<int> cols;
            double lo = 1e300, hi = -1e300;
            for (int a = m.csr_row_ptr[i]; a < m.csr_row_ptr[i + 1]; ++a) {
                const int k = m.csr_col_ind[a];
                if (k == i) continue;
                cols.insert(cols.end(), Pn.ci.begin() + Pn.rp[k], Pn.ci.begin() + Pn.rp[k + 1]);
                const double sk = row_sum(k);
                lo = std::min(lo, sk);
                hi = std::max(hi, sk);
            }
            bool good = Pn.rp[i + 1] > Pn.rp[i];
            for (int a = Pn.rp[i]; a < Pn.rp[i + 1] && good; ++a)
                good = std::find(cols.begin(), cols.end(), Pn.ci[a]) != cols.end();
            const double si = row_sum(i), slack = 1e-12 * std::max(1.0, std::fabs(si));
            if (!(good && si >= lo - slack && si <= hi + slack)) ++new_bad;
        }
        std::printf("  oracle remap invariants: kept rows changed %d, new rows outside their neighbours %d\n",
                    kept_bad, new_bad);
        expect(kept_bad == 0 && new_bad == 0, "oracle remap: kept rows kept, new rows means of neighbours");
    }
    const auto& L = *amg.levels[0];
    std::vector<int> pat_g = down_int(L.P.rp, n + 1), ci_g = down_int(L.P.ci, L.P.nnz);
    pat_g.insert(pat_g.end(), ci_g.begin(), ci_g.end());
    std::vector<int> pat_o = Pn.rp;
    pat_o.insert(pat_o.end(), Pn.ci.begin(), Pn.ci.end());
    same_ints("remapped P pattern", pat_g, pat_o);
    if (pat_g == pat_o) close_rel("remapped P values", down(L.P.v, L.P.nnz), Pn.v, 1e-14);
    const oracle::Csr Rn = oracle::csr_transpose(Pn);
    std::vector<int> rpat_g = down_int(L.R.rp, L.R.rows + 1), rci_g = down_int(L.R.ci, L.R.nnz);
    rpat_g.insert(rpat_g.end(), rci_g.begin(), rci_g.end());
    std::vector<int> rpat_o = Rn.rp;
    rpat_o.insert(rpat_o.end(), Rn.ci.begin(), Rn.ci.end());
    same_ints("R = P^T pattern", rpat_g, rpat_o);
    if (rpat_g == rpat_o) close_rel("R = P^T values", down(L.R.v, L.R.nnz), Rn.v, 1e-14);
    Vec b(n);
    for (int i = 0; i < n; ++i) b[i] = std::cos(0.11 * i);
    DevArray<double> d_b, d_x;
    d_b.upload(b);
    d_x.upload(Vec(n, 0.0));
    GpuLinearSolver solver;
    std::function<void(const double*, double*)> prec = [&amg](const double* r, double* zz) { amg.apply(r, zz); };
    const int it = solver.solve_pcg(n, d_rp, d_ci, d_a, d_b, d_x, 400, 1.0e-10, nullptr, &prec);
    std::printf("  PCG with the remapped hierarchy: %d iterations\n", it);
    expect(solver.converged, "PCG with the remapped hierarchy converges");
}

void case_topology(Fixture& f) {
    std::printf("[mesh topology on the device]\n");
    topology_pair("uniform mesh", f.mesh, true);
    // locally refined: elements with their centroid within 0.08 of (0.35, 0.2), three levels up
    Mesh2D m = f.mesh;
    Vec z(m.num_nodes, 0.0), u = z, v = z, p = z, uo = z, vo = z;
    AdaptiveMesh amr(m, z, z, z);
    AmrParams prm;
    prm.max_level = 5;
    prm.min_level = 0;
    for (int pass = 0; pass < 3; ++pass) {
        amr.extract(m, u, v, p, uo, vo);
        Vec eta(2 * m.elems.size(), 0.0);
        for (size_t e = 0; e < m.elems.size(); ++e) {
            const Triangle& t = m.elems[e];
            const double cx = (m.x[t.v[0]] + m.x[t.v[1]] + m.x[t.v[2]]) / 3.0;
            const double cy = (m.y[t.v[0]] + m.y[t.v[1]] + m.y[t.v[2]]) / 3.0;
            if (std::hypot(cx - 0.35, cy - 0.2) < 0.08) eta[e] = 1.0e3;
        }
        amr.adapt(eta, prm);
    }
    amr.extract(m, u, v, p, uo, vo);
    topology_pair("locally refined mesh", m, true);
    Mesh2D flipped = m;
    std::swap(flipped.elems[flipped.elems.size() / 2].v[1], flipped.elems[flipped.elems.size() / 2].v[2]);
    topology_pair("planted: one element flipped", flipped, false);
    Mesh2D holed = m;
    holed.elems.erase(holed.elems.begin() + (long)(holed.elems.size() / 3));
    topology_pair("planted: one element removed", holed, false);
}

void case_pcg(Fixture& f) {
    std::printf("[PCG with OSS and SA-AMG]\n");
    const PressureSystem o = case_pressure(f, false);
    AmgPreconditioner amg;
    amg.setup(f.n, f.nnz, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val);
    oracle::Amg oa;
    oa.setup(host_csr(f, o.Kp));
    const double tol = 1e-10;
    for (int warm = 0; warm < 2; ++warm) {
        Vec x0(f.n, 0.0);
        if (warm)
            for (int i = 0; i < f.n; ++i) x0[i] = f.tag[i] == 2 ? 0.0 : 0.01 * std::sin(5.0 * f.x[i]);
        f.g.d_dp.upload(x0);
        GpuLinearSolver solver;
        const cfa_oss_op_t op = f.g.oss_op();
        const int it = solver.solve_pcg_fused(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_dp,
                                              4000, tol, &op, &amg);
        Vec x = x0;
        oracle::Op oss_op = [&](const double* in, double* out) {
            oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, o.oss.data(), o.inv_m.data(), f.tag, in, out);
        };
        oracle::Op prec = [&](const double* r, double* z) { oa.apply(r, z); };
        const oracle::SolveInfo si = oracle::pcg(f.n, f.rp, f.ci, o.Kp.data(), o.rhs.data(), x.data(), 4000, tol,
                                                 &oss_op, &prec);
        std::printf("  iterations gpu %d oracle %d (warm start %d)\n", it, si.iters, warm);
        expect(std::abs(it - si.iters) <= 1 && solver.converged && si.converged, "iterations and convergence");
        close_rel("solution", down(f.g.d_dp, f.n), x, 1e-8);
        // the graph-replayed solve and the kernel-by-kernel solve agree
    }
}

// Mixed precision: the FP32 V-cycle (FP64 coarse solve) against the oracle's FP64 cycle, and PCG
// with it against the oracle's FP64 solve (the outer Krylov iteration stays FP64).
void case_mixed(Fixture& f) {
    std::printf("[FP32 SA-AMG cycle inside FP64 PCG]\n");
    const PressureSystem o = case_pressure(f, false);
    AmgPreconditioner amg;
    amg.fp32 = true;
    amg.setup(f.n, f.nnz, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val);
    oracle::Amg oa;
    oa.setup(host_csr(f, o.Kp));
    Vec r(f.n);
    for (auto& v : r) v = std::uniform_real_distribution<double>(-1, 1)(f.rng);
    DevArray<double> d_r, d_z;
    d_r.upload(r);
    d_z.reserve(f.n);
    amg.apply(d_r, d_z);
    Vec z(f.n);
    oa.apply(r.data(), z.data());
    close_rel("one V-cycle, FP32 vs FP64", down(d_z, f.n), z, 1e-5);
    const double tol = 1e-10;
    f.g.d_dp.upload(Vec(f.n, 0.0));
    GpuLinearSolver solver;
    const cfa_oss_op_t op = f.g.oss_op();
    const int it = solver.solve_pcg_fused(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_dp, 4000,
                                          tol, &op, &amg);
    Vec x(f.n, 0.0);
    oracle::Op oss_op = [&](const double* in, double* out) {
        oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, o.oss.data(), o.inv_m.data(), f.tag, in, out);
    };
    oracle::Op prec = [&](const double* rr, double* zz) { oa.apply(rr, zz); };
    const oracle::SolveInfo si = oracle::pcg(f.n, f.rp, f.ci, o.Kp.data(), o.rhs.data(), x.data(), 4000, tol, &oss_op,
                                             &prec);
    std::printf("  iterations FP32 cycle %d, oracle FP64 cycle %d\n", it, si.iters);
    expect(solver.converged && it <= si.iters + si.iters / 4 + 2, "converged, iterations within 25% + 2");
    close_rel("solution", down(f.g.d_dp, f.n), x, 1e-8);
}

void case_bicgstab(Fixture& f) {
    std::printf("[BiCGSTAB2 with ILU(0), FP64 and FP32 factors]\n");
    case_momentum(f, false);   // leaves the momentum system on the device
    const Vec A = down(f.g.d_A_val, f.nnz), bu = down(f.g.d_rhs_u, f.n), bv = down(f.g.d_rhs_v, f.n);
    const Vec u0 = down(f.g.d_u, f.n), v0 = down(f.g.d_v, f.n);
    oracle::Ilu0 o;
    o.set_pattern(f.n, f.rp, f.ci);
    o.factor(A.data());
    Vec ou = u0, ov = v0;
    const oracle::Bicg2Info bi = oracle::bicgstab2(f.n, f.rp, f.ci, A.data(), bu.data(), bv.data(), ou.data(),
                                                   ov.data(), 150, 1e-10, o);
    MulticolourIlu0 ilu;
    ilu.keep_fp32 = true;
    ilu.set_pattern(f.n, f.mesh.csr_row_ptr, f.mesh.csr_col_ind, f.g.d_row_ptr, f.g.d_col_ind);
    ilu.factor(f.g.d_A_val);
    for (int prec = 0; prec < 2; ++prec) {
        f.g.d_u.upload(u0);
        f.g.d_v.upload(v0);
        GpuBicgstab2 sv;
        sv.precond_fp32 = prec == 1;
        sv.solve(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_A_val, f.g.d_rhs_u, f.g.d_rhs_v, f.g.d_u, f.g.d_v, 150, 1e-10,
                 ilu);
        std::printf("  %s factors: iterations gpu (%d, %d) oracle (%d, %d)\n", prec ? "FP32" : "FP64", sv.iters[0],
                    sv.iters[1], bi.iters[0], bi.iters[1]);
        if (!prec)
            expect(std::abs(sv.iters[0] - bi.iters[0]) <= 1 && std::abs(sv.iters[1] - bi.iters[1]) <= 1,
                   "iterations (FP64 factors)");
        expect(sv.converged[0] && sv.converged[1], "converged");
        close_rel(prec ? "u (FP32 factors)" : "u (FP64 factors)", down(f.g.d_u, f.n), ou, prec ? 1e-8 : 1e-9);
        close_rel(prec ? "v (FP32 factors)" : "v (FP64 factors)", down(f.g.d_v, f.n), ov, prec ? 1e-8 : 1e-9);
    }
    // the same solves with the sweeps on the split rows' sliced-ELL copies (built at this size on
    // request; by default from 2^20 rows): iterates bit for bit, FP64 and FP32 factors
    MulticolourIlu0 ilu_csr, ilu_sell;
    ilu_csr.keep_fp32 = ilu_sell.keep_fp32 = true;
    ilu_csr.sell_min_rows = 1 << 30;
    ilu_sell.sell_min_rows = 0;
    ilu_csr.set_pattern(f.n, f.mesh.csr_row_ptr, f.mesh.csr_col_ind, f.g.d_row_ptr, f.g.d_col_ind);
    ilu_sell.set_pattern(f.n, f.mesh.csr_row_ptr, f.mesh.csr_col_ind, f.g.d_row_ptr, f.g.d_col_ind);
    expect(!ilu_csr.has_sell && ilu_sell.has_sell, "CSR rows / sliced-ELL copies as requested");
    ilu_csr.factor(f.g.d_A_val);
    ilu_sell.factor(f.g.d_A_val);
    for (int prec = 0; prec < 2; ++prec) {
        Vec res[2][2];
        int its[2][2];
        for (int k = 0; k < 2; ++k) {
            f.g.d_u.upload(u0);
            f.g.d_v.upload(v0);
            GpuBicgstab2 sv;
            sv.precond_fp32 = prec == 1;
            sv.solve(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_A_val, f.g.d_rhs_u, f.g.d_rhs_v, f.g.d_u, f.g.d_v, 150,
                     1e-10, k ? ilu_sell : ilu_csr);
            res[k][0] = down(f.g.d_u, f.n);
            res[k][1] = down(f.g.d_v, f.n);
            its[k][0] = sv.iters[0];
            its[k][1] = sv.iters[1];
        }
        expect(its[0][0] == its[1][0] && its[0][1] == its[1][1],
               prec ? "iterations: sliced ELL = CSR (FP32 factors)" : "iterations: sliced ELL = CSR (FP64 factors)");
        same_bits(prec ? "u, v: sliced ELL = CSR (FP32 factors)" : "u, v: sliced ELL = CSR (FP64 factors)",
                  [&] { Vec a = res[1][0]; a.insert(a.end(), res[1][1].begin(), res[1][1].end()); return a; }(),
                  [&] { Vec a = res[0][0]; a.insert(a.end(), res[0][1].begin(), res[0][1].end()); return a; }());
    }
}

// The non-default solver paths: Jacobi-preconditioned BiCGSTAB on one momentum component
// (--momentum-precond jacobi, src/cfa/linalg.cuh:261-347) and Jacobi PCG on the PPE with the OSS
// operator (--pressure-precond jacobi, :172-256), against the oracle's restatements.
void case_jacobi(Fixture& f) {
    std::printf("[Jacobi BiCGSTAB and Jacobi PCG]\n");
    case_momentum(f, false);   // leaves the momentum system on the device
    const Vec A = down(f.g.d_A_val, f.nnz), bu = down(f.g.d_rhs_u, f.n), bv = down(f.g.d_rhs_v, f.n);
    const Vec u0 = down(f.g.d_u, f.n), v0 = down(f.g.d_v, f.n);
    const double tol = 1e-10;
    for (int comp = 0; comp < 2; ++comp) {
        const Vec& b = comp ? bv : bu;
        const Vec& x0 = comp ? v0 : u0;
        Vec xo = x0;
        const oracle::SolveInfo si = oracle::bicgstab_jacobi(f.n, f.rp, f.ci, A.data(), b.data(), xo.data(), 200, tol);
        DevArray<double> d_b, d_x;
        d_b.upload(b);
        d_x.upload(x0);
        GpuLinearSolver solver;
        const int it = solver.solve_bicgstab(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_A_val, d_b, d_x, 200, tol);
        std::printf("  component %c: iterations gpu %d oracle %d\n", comp ? 'v' : 'u', it, si.iters);
        expect(std::abs(it - si.iters) <= 1 && solver.converged && si.converged, "iterations and convergence");
        close_rel(comp ? "v (Jacobi BiCGSTAB)" : "u (Jacobi BiCGSTAB)", down(d_x, f.n), xo, 1e-9);
    }
    const PressureSystem o = case_pressure(f, false);
    Vec x(f.n, 0.0);
    f.g.d_dp.upload(x);
    GpuLinearSolver solver;
    const cfa_oss_op_t op = f.g.oss_op();
    const int it = solver.solve_pcg_fused(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_dp, 4000,
                                          tol, &op, nullptr);
    Vec dinv(f.n, 1.0);
    for (int i = 0; i < f.n; ++i)
        for (int k = f.rp[i]; k < f.rp[i + 1]; ++k)
            if (f.ci[k] == i) { dinv[i] = std::fabs(o.Kp[k]) > 1.0e-14 ? 1.0 / o.Kp[k] : 1.0; break; }
    oracle::Op oss_op = [&](const double* in, double* out) {
        oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, o.oss.data(), o.inv_m.data(), f.tag, in, out);
    };
    oracle::Op prec = [&](const double* r, double* z) { for (int i = 0; i < f.n; ++i) z[i] = dinv[i] * r[i]; };
    const oracle::SolveInfo si = oracle::pcg(f.n, f.rp, f.ci, o.Kp.data(), o.rhs.data(), x.data(), 4000, tol, &oss_op,
                                             &prec);
    std::printf("  Jacobi PCG: iterations gpu %d oracle %d\n", it, si.iters);
    // hundreds of iterations: the reduction order moves the stopping step by a few
    expect(solver.converged && si.converged && std::abs(it - si.iters) <= 2 + si.iters / 50,
           "iterations (within 2 + 2%) and convergence");
    close_rel("solution (Jacobi PCG)", down(f.g.d_dp, f.n), x, 1e-8);
}

}  // namespace

// --outlet dirichlet: the upload's pinned node, flux factor and balanced velocities, then the
// momentum rows, the mean-free PPE, the OSS operator, the fused PCG and the correction with the
// outlet parabola, each against the oracle's restatement of the reference's outlet branches
void case_outlet(Fixture& f) {
    std::printf("[outlet dirichlet]\n");
    const Vec u0 = smooth_field(f.mesh, f.rng, 1.5), v0 = smooth_field(f.mesh, f.rng, 0.5);
    const Vec uo0 = smooth_field(f.mesh, f.rng, 1.5), z(f.n, 0.0);
    f.g.outlet_dirichlet = true;
    f.g.sync_from_host(f.mesh, u0, v0, z, uo0, v0);
    const oracle::Outlet out = oracle::outlet_dirichlet_setup(f.mesh);
    expect(f.g.p_pin == out.pin && f.g.p_pin >= 0 && f.tag[f.g.p_pin] == 2, "pinned node (outlet, nearest mid-height)");
    same_bits("outlet parabola factor", Vec{f.g.outlet_scale}, Vec{out.scale});
    same_bits("u flux-balanced on upload", down(f.g.d_u, f.n), oracle::flux_balanced(f.mesh, u0));
    same_bits("u_old flux-balanced on upload", down(f.g.d_u_old, f.n), oracle::flux_balanced(f.mesh, uo0));
    same_bits("u_prev = balanced u_old", down(f.g.d_u_prev, f.n), oracle::flux_balanced(f.mesh, uo0));
    const cfa_mesh_t* m = f.g.view();

    // momentum: outlet rows carry the scaled parabola
    const Vec un = smooth_field(f.mesh, f.rng, 1.5), vn = smooth_field(f.mesh, f.rng, 0.5);
    const Vec p = smooth_field(f.mesh, f.rng, 2.0);
    const double dt = 1.0e-3, a0 = 1.5 / dt, a1 = 2.0 / dt, a2 = -0.5 / dt, umax = 1.2345;
    f.g.d_u_old.upload(un);
    f.g.d_v_old.upload(vn);
    f.g.d_u_prev.upload(u0);
    f.g.d_v_prev.upload(v0);
    f.g.d_p.upload(p);
    f.g.d_u.upload(un);
    f.g.d_v.upload(vn);
    GPU_CHECK(cfa_assemble_momentum(m, f.g.d_u_old, f.g.d_v_old, f.g.d_u_prev, f.g.d_v_prev, f.g.d_p, Sim::NU,
                                    Sim::RHO, a0, a1, a2, 2.0, -1.0, 1.0e-3, 0.5, nullptr, nullptr, umax, -1.0,
                                    f.g.d_A_val, f.g.d_rhs_u, f.g.d_rhs_v, f.g.d_mass_lumped, f.g.d_u, f.g.d_v, f.s()));
    Vec A(f.nnz, 0.0), ru(f.n, 0.0), rv(f.n, 0.0), mass(f.n, 0.0), u = un, v = vn;
    oracle::assemble_momentum(f.ne, f.ev.data(), f.x, f.y, un.data(), vn.data(), u0.data(), v0.data(), p.data(),
                              Sim::NU, Sim::RHO, a0, a1, a2, 2.0, -1.0, 1.0e-3, 0.5, f.rp, f.ci, A.data(), ru.data(),
                              rv.data(), mass.data(), nullptr, nullptr);
    oracle::apply_bc_momentum(f.n, f.tag, f.x, f.y, umax, -1.0, f.rp, f.ci, A.data(), ru.data(), rv.data(), u.data(),
                              v.data(), out);
    same_bits("A (outlet rows identity)", down(f.g.d_A_val, f.nnz), A);
    same_bits("rhs_u", down(f.g.d_rhs_u, f.n), ru);
    same_bits("rhs_v", down(f.g.d_rhs_v, f.n), rv);
    same_bits("u (outlet parabola)", down(f.g.d_u, f.n), u);
    same_bits("v", down(f.g.d_v, f.n), v);

    // pure-Neumann PPE: assembled, made mean-free, pinned
    const Vec us = smooth_field(f.mesh, f.rng, 1.5), vs = smooth_field(f.mesh, f.rng, 0.5);
    const double dtp = (2.0 / 3.0) * 1.0e-3, stab = 0.3;
    f.g.d_u.upload(us);
    f.g.d_v.upload(vs);
    const Vec dp_in = smooth_field(f.mesh, f.rng, 0.1);
    f.g.d_dp.upload(dp_in);
    GPU_CHECK(cfa_lumped_mass(m, f.g.d_mass_lumped, f.s()));
    GPU_CHECK(cfa_oss_weights(m, f.g.d_u, f.g.d_v, dtp, Sim::RHO, Sim::NU, stab, f.g.d_elem_s, f.s()));
    GPU_CHECK(cfa_scalar_gradient(m, f.g.d_p, f.g.d_elem_s, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.s()));
    GPU_CHECK(cfa_assemble_pressure(m, f.g.d_u, f.g.d_v, f.g.d_p, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.g.d_mass_lumped,
                                    f.g.d_elem_s, dtp, Sim::RHO, stab, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_div_num,
                                    f.g.d_dp, f.g.d_oss_val, f.g.d_inv_m, f.s()));
    f.g.pressure_outlet_bc(f.g.d_mass_lumped, f.g.d_rhs_p, f.s());
    Vec pm(f.n, 0.0), es(f.ne, 0.0), gx(f.n, 0.0), gy(f.n, 0.0), div(f.n, 0.0), Kp(f.nnz, 0.0), rhs(f.n, 0.0);
    Vec dp = dp_in, oss(4 * (size_t)f.nnz, 0.0), inv_m(f.n);
    oracle::assemble_lumped_mass(f.ne, f.ev.data(), f.x, f.y, pm.data());
    oracle::compute_oss_weights(f.ne, f.ev.data(), f.x, f.y, us.data(), vs.data(), dtp, Sim::RHO, Sim::NU, stab,
                                es.data());
    oracle::accumulate_scalar_gradient(f.ne, f.ev.data(), f.x, f.y, p.data(), es.data(), gx.data(), gy.data());
    oracle::assemble_pressure_poisson(f.ne, f.ev.data(), f.x, f.y, us.data(), vs.data(), p.data(), gx.data(),
                                      gy.data(), pm.data(), es.data(), dtp, Sim::RHO, stab, f.rp, f.ci, Kp.data(),
                                      rhs.data(), div.data());
    double rhs_abs = 0.0;
    for (double r : rhs) rhs_abs += std::fabs(r);
    oracle::subtract_mean(f.n, rhs.data(), pm.data());
    double rhs_sum = 0.0;
    for (double r : rhs) rhs_sum += r;
    std::printf("  sum rhs after the mean removal %.3e (sum |rhs| %.3e)\n", rhs_sum, rhs_abs);
    expect(std::fabs(rhs_sum) <= 1e-12 * rhs_abs, "oracle rhs mean-free before the pin");
    oracle::apply_bc_pressure(f.n, f.tag, f.rp, f.ci, Kp.data(), rhs.data(), dp.data(), out);
    same_bits("Kp (pinned row and column)", down(f.g.d_Kp_val, f.nnz), Kp);
    same_bits("rhs_p (mean-free, pinned)", down(f.g.d_rhs_p, f.n), rhs);
    same_bits("dp (0 at the pin only)", down(f.g.d_dp, f.n), dp);
    oracle::assemble_oss_operator(f.ne, f.ev.data(), f.x, f.y, es.data(), f.rp, f.ci, f.nnz, oss.data());
    for (int i = 0; i < f.n; ++i) inv_m[i] = 1.0 / std::fmax(pm[i], 1.0e-14);

    // OSS operator: the pinned row is skipped, outlet rows are not
    const Vec in = smooth_field(f.mesh, f.rng, 1.0);
    DevArray<double> d_in, d_out;
    d_in.upload(in);
    d_out.upload(Vec(f.n, 0.0));
    GPU_CHECK(cfa_oss_apply(m, f.g.d_oss_val, f.g.d_inv_m, d_in, d_out, f.g.d_oss_gx, f.g.d_oss_gy, f.s()));
    Vec oo(f.n, 0.0);
    oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, oss.data(), inv_m.data(), f.tag, in.data(), oo.data(), out);
    close_rel("-B^T M^-1 B in (assembled)", down(d_out, f.n), oo, 1e-13);
    d_out.upload(Vec(f.n, 0.0));
    GPU_CHECK(cfa_oss_apply_matrix_free(m, f.g.d_elem_s, f.g.d_mass_lumped, d_in, d_out, f.g.d_oss_gx, f.g.d_oss_gy,
                                        f.s()));
    Vec mf(f.n, 0.0);
    oracle::oss_apply_matrix_free(f.ne, f.n, f.ev.data(), f.x, f.y, f.tag, es.data(), pm.data(), in.data(), mf.data(),
                                  out);
    close_rel("-B^T M^-1 B in (matrix-free)", down(d_out, f.n), mf, 1e-13);

    // fused PCG with SA-AMG on the pinned system
    AmgPreconditioner amg;
    amg.setup(f.n, f.nnz, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val);
    oracle::Amg oa;
    oa.setup(host_csr(f, Kp));
    const double tol = 1e-10;
    f.g.d_dp.upload(z);
    GpuLinearSolver solver;
    const cfa_oss_op_t op = f.g.oss_op();
    const int it = solver.solve_pcg_fused(f.n, f.g.d_row_ptr, f.g.d_col_ind, f.g.d_Kp_val, f.g.d_rhs_p, f.g.d_dp, 4000,
                                          tol, &op, &amg);
    Vec x(f.n, 0.0);
    oracle::Op oss_op = [&](const double* a, double* b) {
        oracle::oss_apply(f.n, f.rp, f.ci, f.nnz, oss.data(), inv_m.data(), f.tag, a, b, out);
    };
    oracle::Op prec = [&](const double* r, double* zz) { oa.apply(r, zz); };
    const oracle::SolveInfo si = oracle::pcg(f.n, f.rp, f.ci, Kp.data(), rhs.data(), x.data(), 4000, tol, &oss_op,
                                             &prec);
    std::printf("  iterations gpu %d oracle %d\n", it, si.iters);
    expect(std::abs(it - si.iters) <= 1 && solver.converged && si.converged, "PCG iterations and convergence");
    close_rel("PCG solution", down(f.g.d_dp, f.n), x, 1e-8);

    // correction: outlet velocities are Dirichlet data and stay
    const Vec dpc = smooth_field(f.mesh, f.rng, 0.1), uc = smooth_field(f.mesh, f.rng, 1.5),
              vc = smooth_field(f.mesh, f.rng, 0.5), pc = smooth_field(f.mesh, f.rng, 2.0);
    f.g.d_dp.upload(dpc);
    f.g.d_u.upload(uc);
    f.g.d_v.upload(vc);
    f.g.d_p.upload(pc);
    f.g.d_mass_lumped.upload(pm);
    f.g.d_div_num.upload(div);
    GPU_CHECK(cfa_correction(m, dtp, Sim::RHO, 1.0, Sim::RHO * Sim::NU, f.g.d_div_num, f.g.d_mass_lumped, f.g.d_dp,
                             f.g.d_u, f.g.d_v, f.g.d_p, f.g.d_vort, f.g.d_grad_dp_x, f.g.d_grad_dp_y, f.g.d_vort_num,
                             f.s()));
    Vec cgx(f.n, 0.0), cgy(f.n, 0.0), cvn(f.n, 0.0), uu = uc, vv = vc, pp = pc, vort(f.n, 0.0);
    oracle::accumulate_grad_dp_and_vorticity(f.ne, f.ev.data(), f.x, f.y, dpc.data(), uc.data(), vc.data(), cgx.data(),
                                             cgy.data(), cvn.data());
    oracle::velocity_pressure_correction(f.n, f.tag, dtp, Sim::RHO, 1.0, Sim::RHO * Sim::NU, div.data(), pm.data(),
                                         cgx.data(), cgy.data(), cvn.data(), dpc.data(), uu.data(), vv.data(),
                                         pp.data(), vort.data(), out);
    same_bits("u corrected (outlet untouched)", down(f.g.d_u, f.n), uu);
    same_bits("v corrected", down(f.g.d_v, f.n), vv);
    same_bits("p updated", down(f.g.d_p, f.n), pp);

    f.g.outlet_dirichlet = false;   // back to the default outlet for later cases
    f.g.sync_from_host(f.mesh, z, z, z, z, z);
}

int main(int argc, char** argv) {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);   // each line reaches a pipe at once
    const std::string what = argc > 1 ? argv[1] : "all";
    const int level = argc > 2 ? std::atoi(argv[2]) : 2;
    GPU_CHECK(cfa_init(0));
    int status = 0;
    {
        Fixture f(level);
        const bool all = what == "all";
        if (all || what == "momentum") { case_momentum(f, false); case_momentum(f, true); }
        if (all || what == "pressure") case_pressure(f);
        if (all || what == "correction") case_correction(f);
        if (all || what == "forces") case_forces(f);
        if (all || what == "indicator") case_indicator(f);
        if (all || what == "mms") case_mms(f);
        if (all || what == "oss") case_oss(f);
        if (all || what == "ilu") case_ilu(f);
        if (all || what == "spmv") case_spmv(f);
        if (all || what == "amg") case_amg(f);
        if (all || what == "sell") case_sell(f);
        if (all || what == "pcg") case_pcg(f);
        if (all || what == "bicgstab") case_bicgstab(f);
        if (all || what == "jacobi") case_jacobi(f);
        if (all || what == "mixed") case_mixed(f);
        if (all || what == "topology") case_topology(f);
        if (all || what == "outlet") case_outlet(f);
        if (all || what == "nvb") case_nvb(f, level);
        if (all || what == "remap") case_remap(f);
        if (all || what == "reproject") case_reproject(f);
        if (all || what == "noalloc") case_noalloc(f, level);
        std::printf(failures ? "%d FAILURE(S)\n" : "ALL PASSED\n", failures);
        status = failures ? 1 : 0;
    }   // the fixture's device memory is freed here
    // End without the runtimes' exit-time teardown: short runs of this program have ended with
    // SIGSEGV after all checks had passed and printed, i.e. inside that teardown.
    GPU_CHECK(cfa_device_sync());
    std::fflush(stdout);
    std::fflush(stderr);
    std::_Exit(status);
}
