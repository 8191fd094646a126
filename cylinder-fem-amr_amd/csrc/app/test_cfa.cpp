// Unit tests of the host layer and the solver classes, covering the checks of the reference's
// tests/test_cfa.cu: NVB adaptation (conformity, level cap, angle bound, exact transfer of linear
// fields), refine/coarsen round trip, the minimum level, retransfer, the zero force of a fluid at
// rest, the multicolour ILU(0) (valid colouring, the ILU(0) identity on the pattern) and the
// two-component BiCGSTAB against the one-component Jacobi solver, the assembled OSS projection
// against its matrix-free form, and the AMG remap.
//
//   test_cfa [gpu|all]      (default all; every check runs on the GPU)
//
// Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>

#include "../host/device.hpp"

using namespace cfa;

namespace {

int failures = 0;

#define EXPECT(cond, ...)                                                   \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ++failures;                                                     \
            std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);       \
            std::fprintf(stderr, __VA_ARGS__);                              \
            std::fprintf(stderr, "\n");                                     \
        }                                                                   \
    } while (0)

Mesh2D make_base() {
    Mesh2D m;
    generate_cylinder_mesh(m, 8, 7, 28);
    build_csr_and_boundary_topology(m);
    return m;
}

double plane(double x, double y) { return 0.3 + 1.7 * x - 2.9 * y; }

bool near_cylinder(double x, double y) {
    // NVB midpoints on the cylinder are moved onto the arc but keep their chord value; nodes
    // interpolated from them are exempt from the exact-transfer checks
    return std::hypot(x - Sim::X_CYL, y - Sim::Y_CYL) < Sim::R_CYL + 0.03;
}

// An adaptive mesh with a linear pressure field and zero velocities
struct World {
    Mesh2D mesh;
    std::vector<double> u, v, p, uo, vo;
    AdaptiveMesh amr;
    explicit World(const Mesh2D& base)
        : mesh(base), u(base.num_nodes, 0.0), v(u), p(linear_field(base)), uo(u), vo(u), amr(base, u, v, p) {
        amr.extract(mesh, u, v, p, uo, vo);
    }
    static std::vector<double> linear_field(const Mesh2D& m) {
        std::vector<double> r(m.num_nodes);
        for (int i = 0; i < m.num_nodes; ++i) r[i] = plane(m.x[i], m.y[i]);
        return r;
    }
    MeshQuality refresh() {
        amr.extract(mesh, u, v, p, uo, vo);
        build_csr_and_boundary_topology(mesh);
        MeshQuality q;
        EXPECT(validate_mesh(mesh, q), "adapted mesh is not a valid conforming triangulation");
        return q;
    }
    void adapt(const std::vector<double>& eta, const AmrParams& prm) {
        amr.scatter_solution(u, v, p, uo, vo);
        amr.adapt(eta, prm);
        refresh();
    }
    void centroid(int e, double& xc, double& yc) const {
        const Triangle& t = mesh.elems[e];
        xc = (mesh.x[t.v[0]] + mesh.x[t.v[1]] + mesh.x[t.v[2]]) / 3.0;
        yc = (mesh.y[t.v[0]] + mesh.y[t.v[1]] + mesh.y[t.v[2]]) / 3.0;
    }
};

AmrParams free_params(int max_level) {
    AmrParams prm;
    prm.max_level = max_level;
    prm.tol_vort = 1.0;
    prm.tol_interp = 1.0;
    prm.lookahead_time = 0.0;
    prm.keep_radius = 0.0;
    prm.min_level = 0;
    prm.max_elems = 400000;
    return prm;
}

// ---- host-side checks --------------------------------------------------------------------------
void random_adaptation() {
    const Mesh2D base = make_base();
    MeshQuality q0;
    validate_mesh(base, q0);
    World w(base);
    const AmrParams prm = free_params(4);
    std::mt19937 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    for (int cycle = 0; cycle < 12; ++cycle) {
        const int L = w.mesh.num_elems;
        std::vector<double> eta(2 * L);
        const double bx = 0.2 + 1.8 * U(rng), by = 0.41 * U(rng);
        for (int e = 0; e < L; ++e) {
            double xc, yc;
            w.centroid(e, xc, yc);
            const double d2 = (xc - bx) * (xc - bx) + (yc - by) * (yc - by);
            eta[e] = 8.0 * std::exp(-d2 / 0.01) * (0.5 + U(rng));
            eta[L + e] = 0.05 * U(rng);
        }
        w.amr.scatter_solution(w.u, w.v, w.p, w.uo, w.vo);
        w.amr.adapt(eta, prm);
        const MeshQuality q = w.refresh();
        EXPECT(q.max_level <= prm.max_level, "cycle %d: level %d above the cap %d", cycle, q.max_level, prm.max_level);
        EXPECT(q.min_angle_deg >= q0.min_angle_deg - 1.0, "cycle %d: min angle %.2f vs base %.2f", cycle,
               q.min_angle_deg, q0.min_angle_deg);
        EXPECT(std::fabs(q.total_area - q0.total_area) < 1.0e-3 * q0.total_area, "cycle %d: area %.9f vs %.9f",
               cycle, q.total_area, q0.total_area);
        double err = 0.0;
        for (int i = 0; i < w.mesh.num_nodes; ++i)
            if (!near_cylinder(w.mesh.x[i], w.mesh.y[i]))
                err = std::max(err, std::fabs(w.p[i] - plane(w.mesh.x[i], w.mesh.y[i])));
        EXPECT(err < 1.0e-12, "cycle %d: linear field not transferred exactly (%.3e)", cycle, err);
    }
    std::printf("  random adaptation: %d elements after 12 cycles\n", w.mesh.num_elems);
}

void round_trip() {
    const Mesh2D base = make_base();
    World w(base);
    const AmrStats st = w.amr.refine_uniform(2);
    const MeshQuality q = w.refresh();
    EXPECT(w.mesh.num_elems == 16 * base.num_elems, "uniform level 2: %d elements", w.mesh.num_elems);
    EXPECT(st.closure_bisections == 0, "uniform refinement needed %d closure bisections", st.closure_bisections);
    EXPECT(q.max_level == 2, "uniform level 2 reports max level %d", q.max_level);
    AmrParams prm = free_params(4);
    prm.tol_vort = 0.1;
    prm.tol_interp = 0.01;
    prm.max_elems = 200000;
    for (int cycle = 0; cycle < 6; ++cycle) w.adapt(std::vector<double>(2 * w.mesh.num_elems, 0.0), prm);
    EXPECT(w.mesh.num_elems == base.num_elems && w.mesh.num_nodes == base.num_nodes,
           "after coarsening: %d elements / %d nodes (base %d / %d)", w.mesh.num_elems, w.mesh.num_nodes,
           base.num_elems, base.num_nodes);
    std::vector<std::pair<double, double>> a, b;
    for (int i = 0; i < base.num_nodes; ++i) a.emplace_back(base.x[i], base.y[i]);
    for (int i = 0; i < w.mesh.num_nodes; ++i) b.emplace_back(w.mesh.x[i], w.mesh.y[i]);
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    EXPECT(a == b, "coarsened mesh vertices differ from the base mesh");
}

void min_level() {
    const Mesh2D base = make_base();
    World w(base);
    w.amr.refine_uniform(3);
    w.refresh();
    AmrParams prm = free_params(4);
    prm.tol_vort = 0.1;
    prm.tol_interp = 0.01;
    prm.min_level = 2;
    for (int cycle = 0; cycle < 6; ++cycle) w.adapt(std::vector<double>(2 * w.mesh.num_elems, 0.0), prm);
    int below = 0;
    for (const Triangle& t : w.mesh.elems) below += t.level < prm.min_level ? 1 : 0;
    EXPECT(below == 0 && w.mesh.num_elems == 16 * base.num_elems, "min level 2: %d below, %d elements", below,
           w.mesh.num_elems);
}

void retransfer() {
    World w(make_base());
    w.amr.refine_uniform(2);
    w.refresh();
    const AmrParams prm = free_params(4);
    std::mt19937 rng(99);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto later = [](double x, double y) { return -1.1 + 0.4 * x + 3.3 * y; };
    double same = 0.0, lin = 0.0;
    for (int cycle = 0; cycle < 4; ++cycle) {
        const Mesh2D& m = w.mesh;
        for (int i = 0; i < m.num_nodes; ++i) {
            w.u[i] = std::sin(3.0 * m.x[i]) * m.y[i];
            w.v[i] = std::cos(2.0 * m.y[i]);
            w.uo[i] = 0.9 * w.u[i];
            w.vo[i] = 0.9 * w.v[i];
            w.p[i] = m.x[i] * m.x[i] - m.y[i];
        }
        const std::vector<double> su = w.u, sv = w.v, sp = w.p, suo = w.uo, svo = w.vo;
        std::vector<double> lp(m.num_nodes);
        for (int i = 0; i < m.num_nodes; ++i) lp[i] = later(m.x[i], m.y[i]);
        const std::vector<int> old_vids = w.amr.node_vertices();
        const int L = m.num_elems;
        std::vector<double> eta(2 * L);
        const double bx = 0.3 + 1.5 * U(rng), by = 0.41 * U(rng);
        for (int e = 0; e < L; ++e) {
            double xc, yc;
            w.centroid(e, xc, yc);
            eta[e] = 8.0 * std::exp(-((xc - bx) * (xc - bx) + (yc - by) * (yc - by)) / 0.01);
            eta[L + e] = 0.05 * U(rng);
        }
        w.amr.scatter_solution(w.u, w.v, w.p, w.uo, w.vo);
        w.amr.adapt(eta, prm);
        w.refresh();
        std::vector<double> ru, rv, rp, ruo, rvo;
        w.amr.retransfer(old_vids, su, sv, sp, suo, svo, ru, rv, rp, ruo, rvo);
        for (int i = 0; i < w.mesh.num_nodes; ++i)
            same = std::max({same, std::fabs(ru[i] - w.u[i]), std::fabs(rv[i] - w.v[i]), std::fabs(rp[i] - w.p[i]),
                             std::fabs(ruo[i] - w.uo[i]), std::fabs(rvo[i] - w.vo[i])});
        w.amr.retransfer(old_vids, su, sv, lp, suo, svo, ru, rv, rp, ruo, rvo);
        for (int i = 0; i < w.mesh.num_nodes; ++i)
            if (!near_cylinder(w.mesh.x[i], w.mesh.y[i]))
                lin = std::max(lin, std::fabs(rp[i] - later(w.mesh.x[i], w.mesh.y[i])));
    }
    EXPECT(same == 0.0, "retransfer of the same state differs from adapt()'s transfer: %.3e", same);
    EXPECT(lin < 1.0e-12, "retransfer of a later linear field not exact: %.3e", lin);
    std::printf("  retransfer: 4 cycles, %d elements at the end\n", w.mesh.num_elems);
}

// ---- device checks -----------------------------------------------------------------------------
void force_at_rest() {
    Mesh2D mesh = make_base();
    std::vector<double> z(mesh.num_nodes, 0.0), p(mesh.num_nodes, 7.25);
    DeviceMeshBuffers g;
    g.sync_from_host(mesh, z, z, p, z, z);
    const int probe[4] = {g.probe_front.a, g.probe_front.b, g.probe_back.a, g.probe_back.b};
    GPU_CHECK(cfa_forces(g.view(), g.d_cyl_edges, g.n_cyl_edges, g.d_layer, g.n_layer, g.d_u, g.d_v, g.d_u_old,
                         g.d_v_old, g.d_u_prev, g.d_v_prev, g.d_p, 1500.0, -2000.0, 500.0, Sim::NU, Sim::RHO, probe,
                         g.probe_front.w, g.probe_back.w, g.d_diag, g.d_ws, solver_stream()));
    double d[6];
    g.d_diag.download(d, 6);
    EXPECT(std::fabs(d[0]) < 1.0e-12 && std::fabs(d[1]) < 1.0e-12, "force at rest: (%.3e, %.3e)", d[0], d[1]);
    EXPECT(std::fabs(d[2] - 7.25) < 1.0e-12 && std::fabs(d[3] - 7.25) < 1.0e-12, "probes: %.15g %.15g", d[2], d[3]);
}

void momentum_ilu() {
    World w(make_base());
    w.amr.refine_uniform(2);
    w.refresh();
    const Mesh2D& m = w.mesh;
    const int n = m.num_nodes, nnz = m.nnz;
    const std::vector<int>& rp = m.csr_row_ptr;
    const std::vector<int>& ci = m.csr_col_ind;
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> a(nnz), b(2 * n), x0(2 * n, 0.0);
    for (int i = 0; i < n; ++i) {   // non-symmetric, diagonally dominant (convection-like)
        double off = 0.0;
        int d = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] == i) { d = k; continue; }
            a[k] = -(0.2 + U(rng)) * (ci[k] > i ? 1.6 : 0.6);
            off += std::fabs(a[k]);
        }
        a[d] = 1.05 * off + 0.1;
        b[i] = U(rng) - 0.5;
        b[n + i] = std::sin(0.01 * i);
    }
    DevArray<int> d_rp, d_ci;
    DevArray<double> d_a, d_b, d_x, d_xj;
    d_rp.upload(rp);
    d_ci.upload(ci);
    d_a.upload(a);
    d_b.upload(b);
    d_x.upload(x0);
    d_xj.upload(x0);
    MulticolourIlu0 ilu;
    ilu.set_pattern(n, rp, ci, d_rp, d_ci);
    ilu.factor(d_a);
    const std::vector<int> col = ilu.d_colour.to_host(n);
    const std::vector<double> lu = ilu.d_lu.to_host(nnz);
    int clash = 0;
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) clash += (ci[k] != i && col[ci[k]] == col[i]) ? 1 : 0;
    EXPECT(clash == 0, "colouring: %d neighbouring pairs share a colour", clash);
    // (LU)_ij = sum_k l_ik u_kj over k in both patterns with k <= min(i, j) in (colour, index) order
    auto pos = [&](int r, int c) {
        for (int k = rp[r]; k < rp[r + 1]; ++k)
            if (ci[k] == c) return k;
        return -1;
    };
    auto earlier = [&](int p, int q) { return col[p] < col[q] || (col[p] == col[q] && p < q); };
    double worst = 0.0;
    for (int i = 0; i < n; ++i)
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const int j = ci[e];
            double s = 0.0;
            for (int q = rp[i]; q < rp[i + 1]; ++q) {
                const int k = ci[q];
                if (!(k == i || earlier(k, i)) || !(k == j || earlier(k, j))) continue;
                const int kj = pos(k, j);
                if (kj >= 0) s += (k == i ? 1.0 : lu[q]) * lu[kj];
            }
            worst = std::max(worst, std::fabs(s - a[e]) / std::fabs(a[pos(i, i)]));
        }
    EXPECT(worst < 1.0e-12, "ILU(0) identity: max |(LU)_ij - a_ij| / a_ii = %.3e", worst);
    GpuBicgstab2 two;
    two.solve(n, d_rp, d_ci, d_a, d_b, d_b + n, d_x, d_x + n, 200, 1.0e-10, ilu);
    GpuLinearSolver one;
    one.solve_bicgstab(n, d_rp, d_ci, d_a, d_b, d_xj, 400, 1.0e-10);
    one.solve_bicgstab(n, d_rp, d_ci, d_a, d_b + n, d_xj + n, 400, 1.0e-10);
    const std::vector<double> x = d_x.to_host(2 * n), xj = d_xj.to_host(2 * n);
    double diff = 0.0, scale = 0.0;
    for (int i = 0; i < 2 * n; ++i) {
        diff = std::max(diff, std::fabs(x[i] - xj[i]));
        scale = std::max(scale, std::fabs(xj[i]));
    }
    EXPECT(two.converged[0] && two.converged[1], "two-component BiCGSTAB did not converge");
    EXPECT(diff < 1.0e-8 * scale, "ILU and Jacobi solutions differ: %.3e (scale %.3e)", diff, scale);
    std::printf("  momentum ILU: %d nodes, %d colours, iterations (%d, %d)\n", n, ilu.n_colours, two.iters[0],
                two.iters[1]);
}

void oss_assembled() {
    World w(make_base());
    w.amr.refine_uniform(2);
    w.refresh();
    const Mesh2D& m = w.mesh;
    std::mt19937 rng(11);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> z(m.num_nodes, 0.0), xin(m.num_nodes), s(m.num_elems);
    for (double& v : xin) v = U(rng) - 0.5;
    for (double& v : s) v = 0.1 + U(rng);
    DeviceMeshBuffers g;
    g.sync_from_host(m, z, z, z, z, z);
    const int n = g.n_nodes;
    g.d_elem_s.upload(s);
    GPU_CHECK(cfa_lumped_mass(g.view(), g.d_mass_lumped, solver_stream()));
    DevArray<double> d_in, d_a, d_b;
    d_in.upload(xin);
    d_a.upload(z);
    d_b.upload(z);
    apply_oss_projection_matrix_free(g, d_in, d_a);
    assemble_oss_projection(g);
    apply_oss_projection(g, d_in, d_b);
    const std::vector<double> a = d_a.to_host(n), b = d_b.to_host(n);
    double diff = 0.0, scale = 0.0;
    for (int i = 0; i < n; ++i) {
        diff = std::max(diff, std::fabs(a[i] - b[i]));
        scale = std::max(scale, std::fabs(a[i]));
    }
    EXPECT(scale > 0.0 && diff < 1.0e-12 * scale, "assembled OSS projection differs: %.3e (scale %.3e)", diff,
           scale);
}

// shifted graph Laplacian of the mesh graph (SPD, nearly singular)
void graph_laplacian(const Mesh2D& m, DevArray<int>& d_rp, DevArray<int>& d_ci, DevArray<double>& d_a) {
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
    d_rp.upload(m.csr_row_ptr);
    d_ci.upload(m.csr_col_ind);
    d_a.upload(a);
}

void amg_remap() {
    auto vcycle = [](AmgPreconditioner& amg, int n) {
        std::vector<double> r(n);
        for (int i = 0; i < n; ++i) r[i] = std::sin(0.37 * i) + 0.2;
        DevArray<double> d_r, d_z;
        d_r.upload(r);
        d_z.reserve(n);
        amg.apply(d_r, d_z);
        return d_z.to_host(n);
    };
    World w(make_base());
    w.amr.refine_uniform(3);
    w.refresh();
    DevArray<int> d_rp, d_ci;
    DevArray<double> d_a;
    graph_laplacian(w.mesh, d_rp, d_ci, d_a);
    const int n0 = w.mesh.num_nodes;
    AmgPreconditioner amg;
    amg.keep_p0_host = true;
    amg.setup(n0, w.mesh.nnz, d_rp, d_ci, d_a);
    EXPECT(amg.num_levels() > 2, "hierarchy too shallow (%d levels)", amg.num_levels());
    const std::vector<double> z_setup = vcycle(amg, n0);
    std::vector<int> ident(n0);
    for (int i = 0; i < n0; ++i) ident[i] = i;
    amg.remap(n0, w.mesh.nnz, d_rp, d_ci, d_a, w.mesh.csr_row_ptr.data(), w.mesh.csr_col_ind.data(), ident);
    EXPECT(amg.remaps == 1, "remap fell back to setup on an unchanged mesh");
    const std::vector<double> z_remap = vcycle(amg, n0);
    double dz = 0.0, sz = 0.0;
    for (int i = 0; i < n0; ++i) {
        dz = std::max(dz, std::fabs(z_remap[i] - z_setup[i]));
        sz = std::max(sz, std::fabs(z_setup[i]));
    }
    EXPECT(dz <= 1.0e-14 * sz, "remap on an unchanged mesh changes the V-cycle: %.3e (scale %.3e)", dz, sz);

    const std::vector<int> old_vids = w.amr.node_vertices();
    const std::vector<double> old_x = w.mesh.x, old_y = w.mesh.y;
    AmrParams prm = free_params(5);
    const int L = w.mesh.num_elems;
    std::vector<double> eta(2 * L, 0.5);
    for (int e = 0; e < L; ++e) {
        double xc, yc;
        w.centroid(e, xc, yc);
        eta[e] = ((xc - 0.4) * (xc - 0.4) + (yc - 0.2) * (yc - 0.2)) < 0.0004 ? 8.0 : 0.5;
    }
    w.adapt(eta, prm);
    const Mesh2D& m = w.mesh;
    const int n = m.num_nodes;
    std::vector<int> slot_old(w.amr.num_vertex_slots(), -1), new_to_old(n);
    for (int j = 0; j < (int)old_vids.size(); ++j) slot_old[old_vids[j]] = j;
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        const int j = slot_old[w.amr.node_vertices()[i]];
        new_to_old[i] = (j >= 0 && old_x[j] == m.x[i] && old_y[j] == m.y[i]) ? j : -1;
        kept += new_to_old[i] >= 0 ? 1 : 0;
    }
    EXPECT(n > n0 && kept < n, "refinement added no nodes (%d -> %d, kept %d)", n0, n, kept);
    graph_laplacian(m, d_rp, d_ci, d_a);
    std::vector<double> b(n);
    for (int i = 0; i < n; ++i) b[i] = std::cos(0.11 * i);
    DevArray<double> d_b, d_x;
    d_b.upload(b);
    d_x.reserve(n);
    std::function<void(const double*, double*)> prec = [&amg](const double* r, double* zz) { amg.apply(r, zz); };
    GpuLinearSolver solver;
    amg.remap(n, m.nnz, d_rp, d_ci, d_a, m.csr_row_ptr.data(), m.csr_col_ind.data(), new_to_old);
    EXPECT(amg.remaps == 2, "remap fell back to setup after a local refinement");
    dev_zero(d_x, n * sizeof(double));
    const int it_remap = solver.solve_pcg(n, d_rp, d_ci, d_a, d_b, d_x, 400, 1.0e-10, nullptr, &prec);
    const bool conv_remap = solver.converged;
    amg.setup(n, m.nnz, d_rp, d_ci, d_a);
    dev_zero(d_x, n * sizeof(double));
    const int it_setup = solver.solve_pcg(n, d_rp, d_ci, d_a, d_b, d_x, 400, 1.0e-10, nullptr, &prec);
    EXPECT(conv_remap && solver.converged, "PCG did not converge (remap %d, setup %d)", (int)conv_remap,
           (int)solver.converged);
    EXPECT(it_remap <= 3 * it_setup + 5, "remapped hierarchy is a poor preconditioner: %d vs %d", it_remap, it_setup);
    std::printf("  AMG remap: %d -> %d nodes (%d new), PCG iterations remap %d, fresh setup %d\n", n0, n, n - kept,
                it_remap, it_setup);
}

// Host check of the ILU(0) pattern's sliced-ELL layout (MulticolourIlu0::build_pattern, no device):
// on a 120 x 120 grid graph with the nine-point stencil, for L and U, every colour's rows sit once
// in their 256-position window, in decreasing length with ties by row, the slice offsets give each
// slice its first row's length, every slot of a row holds its split entry (column and index) in
// order, and padding holds column 0 and index -1.
void ilu_sell_layout() {
    const int side = 120, n = side * side;
    std::vector<int> rp(1, 0), ci;
    for (int j = 0; j < side; ++j)
        for (int i = 0; i < side; ++i) {
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    const int a = i + di, b = j + dj;
                    if (a >= 0 && b >= 0 && a < side && b < side) ci.push_back(b * side + a);
                }
            rp.push_back((int)ci.size());
        }
    const MulticolourIlu0::Pattern P = MulticolourIlu0::build_pattern(n, rp, ci, 0);
    EXPECT(!P.s_pos0.empty(), "no sliced-ELL layout was built");
    if (P.s_pos0.empty()) return;
    const int nc = P.n_colours;
    for (int part = 0; part < 2; ++part) {
        const std::vector<int>& ptr = part ? P.up : P.lp;
        const std::vector<int>& col = part ? P.uc : P.lc;
        const std::vector<int>& row = part ? P.s_urow : P.s_lrow;
        const std::vector<int>& sp = part ? P.s_usp : P.s_lsp;
        const std::vector<int>& sc = part ? P.s_uc : P.s_lc;
        const std::vector<int>& src = part ? P.s_usrc : P.s_lsrc;
        int bad_place = 0, bad_order = 0, bad_len = 0, bad_slot = 0, bad_pad = 0, longest = 0;
        for (int c = 0; c < nc; ++c) {
            const int cs = P.colour_start[c], m = P.colour_start[c + 1] - cs;
            std::vector<int> seen(m, 0);
            for (int p = P.s_pos0[c]; p < P.s_pos0[c + 1]; ++p) {
                const int t = row[p], w0 = P.s_pos0[c] + (p - P.s_pos0[c]) / 256 * 256;
                const int q = p / 64, sl = (sp[q + 1] - sp[q]) / 64;
                if (t < 0) {
                    bad_place += p - P.s_pos0[c] < m ? 1 : 0;
                    for (int k = 0; k < sl; ++k) bad_pad += sc[sp[q] + 64 * k + p % 64] != 0 || src[sp[q] + 64 * k + p % 64] != -1;
                    continue;
                }
                if (t >= m || t < w0 - P.s_pos0[c] || t >= w0 - P.s_pos0[c] + 256 || seen[t]++) {
                    ++bad_place;
                    continue;
                }
                const int i = cs + t, len = ptr[i + 1] - ptr[i];
                longest = std::max(longest, len);
                if (p > w0 && row[p - 1] >= 0) {
                    const int tp = row[p - 1], lp = ptr[cs + tp + 1] - ptr[cs + tp];
                    bad_order += (lp < len || (lp == len && tp > t)) ? 1 : 0;
                }
                bad_len += (sl < len || (p % 64 == 0 && sl != len)) ? 1 : 0;
                for (int k = 0; k < sl; ++k) {
                    const int at = sp[q] + 64 * k + p % 64;
                    if (k < len) bad_slot += sc[at] != col[ptr[i] + k] || src[at] != ptr[i] + k;
                    else bad_pad += sc[at] != 0 || src[at] != -1;
                }
            }
            for (int t = 0; t < m; ++t) bad_place += seen[t] == 1 ? 0 : 1;
        }
        const char* what = part ? "U" : "L";
        EXPECT(bad_place == 0, "%s rows misplaced: %d", what, bad_place);
        EXPECT(bad_order == 0, "%s window order broken: %d", what, bad_order);
        EXPECT(bad_len == 0, "%s slice lengths wrong: %d", what, bad_len);
        EXPECT(bad_slot == 0 && bad_pad == 0, "%s slots wrong: %d, padding wrong: %d", what, bad_slot, bad_pad);
        EXPECT(longest == (part ? P.s_ulen : P.s_llen), "%s longest row %d, recorded %d", what, longest,
               part ? P.s_ulen : P.s_llen);
        std::printf("  ILU sliced ELL (%s): %d colours, %d slots for %d entries, longest row %d\n", what, nc,
                    (int)sc.size(), (int)col.size(), longest);
    }
    const MulticolourIlu0::Pattern Q = MulticolourIlu0::build_pattern(n, rp, ci);
    EXPECT(Q.s_pos0.empty(), "a layout below the default row threshold");
}

}  // namespace

int main(int argc, char** argv) {
    const std::string what = argc > 1 ? argv[1] : "all";
    if (what != "gpu" && what != "host" && what != "all") {
        std::fprintf(stderr, "usage: test_cfa [gpu|host|all] (host: the checks that need no device)\n");
        return 2;
    }
    std::printf("cylinder_fem_amr unit tests (%s)\n", what.c_str());
    if (what != "gpu") ilu_sell_layout();
    if (what != "host") {
        GPU_CHECK(cfa_init(0));
        random_adaptation();
        round_trip();
        min_level();
        retransfer();
        force_at_rest();
        momentum_ilu();
        oss_assembled();
        amg_remap();
    }
    std::printf(failures == 0 ? "ALL PASSED\n" : "%d FAILURE(S)\n", failures);
    return failures == 0 ? 0 : 1;
}
