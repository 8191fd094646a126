// Device buffers and GPU driver routines (src/cfa/device.cuh), over the C ABI.
#include "device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <iomanip>
#include <iostream>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cfa {

std::vector<KernelTime>& kernel_timings() {
    static std::vector<KernelTime> k;
    return k;
}

namespace {
double wrap_angle(double a) {
    while (a <= -Sim::PI) a += 2.0 * Sim::PI;
    while (a > Sim::PI) a -= 2.0 * Sim::PI;
    return a;
}
}  // namespace

// src/cfa/diagnostics.cuh:23-45
PressureProbe find_cylinder_probe(const Mesh2D& mesh, double theta) {
    PressureProbe best;
    double best_dist = 1.0e300;
    for (const auto& ed : mesh.cylinder_edges) {
        const int a = ed.first, b = ed.second;
        const double ta = std::atan2(mesh.y[a] - Sim::Y_CYL, mesh.x[a] - Sim::X_CYL);
        const double tb = std::atan2(mesh.y[b] - Sim::Y_CYL, mesh.x[b] - Sim::X_CYL);
        const double span = wrap_angle(tb - ta), s = wrap_angle(theta - ta);
        if (std::abs(span) < 1.0e-14) continue;
        const double w = s / span;
        if (w >= -1.0e-12 && w <= 1.0 + 1.0e-12) {
            best.a = a;
            best.b = b;
            best.w = std::min(1.0, std::max(0.0, w));
            return best;
        }
        const double dist = std::min(std::abs(s), std::abs(wrap_angle(theta - tb)));
        if (dist < best_dist) {
            best_dist = dist;
            best.a = a;
            best.b = b;
            best.w = (std::abs(s) < std::abs(wrap_angle(theta - tb))) ? 0.0 : 1.0;
        }
    }
    return best;
}

bool device_topology_default() {
    static const bool on = [] {
        const char* e = std::getenv("CFA_HOST_TOPOLOGY");
        return !(e && e[0] == '1');
    }();
    return on;
}

void build_incidence(const Mesh2D& mesh, std::vector<int>& inc_ptr, std::vector<int>& inc_elem,
                     std::vector<int>& inc_pos) {
    const int n = mesh.num_nodes, ne = mesh.num_elems;
    inc_ptr.assign(n + 1, 0);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) inc_ptr[mesh.elems[e].v[k] + 1]++;
    for (int i = 0; i < n; ++i) inc_ptr[i + 1] += inc_ptr[i];
    inc_elem.resize(3 * (size_t)ne);
    {
        std::vector<int> pos(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) inc_elem[pos[mesh.elems[e].v[k]]++] = 3 * e + k;
    }
    inc_pos.resize(9 * (size_t)ne);
    const int* rp = mesh.csr_row_ptr.data();
    const int* ci = mesh.csr_col_ind.data();
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
            const int e = inc_elem[s] / 3;
            for (int b = 0; b < 3; ++b) {
                const int col = mesh.elems[e].v[b];
                const int* it = std::lower_bound(ci + rp[i], ci + rp[i + 1], col);
                inc_pos[3 * (size_t)s + b] = (it != ci + rp[i + 1] && *it == col) ? (int)(it - ci) : -1;
            }
        }
}

cfa_oss_op_t DeviceMeshBuffers::oss_op() const {
    cfa_oss_op_t op;
    op.oss_val = d_oss_val;
    op.inv_m = d_inv_m;
    op.bc_tag = d_bc_tag;
    op.gx = d_oss_gx;
    op.gy = d_oss_gy;
    op.outlet_dirichlet = outlet_dirichlet ? 1 : 0;
    op.p_pin = p_pin;
    return op;
}

void DeviceMeshBuffers::pressure_outlet_bc(const double* mass, double* rhs, cfa_stream_t s) {
    if (!outlet_dirichlet || n_nodes == 0) return;
    d_mean_ws_.reserve(cfa_pressure_mean_ws_doubles(n_nodes), 0);
    GPU_CHECK(cfa_pressure_remove_mean(&view_, mass, rhs, d_mean_ws_, s));
}

namespace {

// src/cfa/device.cuh:153-172: the pinned node (the outlet node nearest mid-height, the first in
// index order on ties) and the outlet parabola factor, inlet over outlet of the trapezoidal flux
// of the unit parabola y (H - y) through the boundary's nodes and the two corners.
void outlet_pin_and_scale(const Mesh2D& mesh, int& pin, double& scale) {
    const double H = Sim::H_CHANNEL;
    const int n = (int)mesh.x.size();
    pin = -1;
    for (int i = 0; i < n; ++i)
        if (mesh.bc_tag[i] == BC_OUTLET &&
            (pin < 0 || std::fabs(mesh.y[i] - 0.5 * H) < std::fabs(mesh.y[pin] - 0.5 * H)))
            pin = i;
    if (pin < 0) {
        std::fprintf(stderr, "--outlet dirichlet: no outlet node\n");
        std::exit(EXIT_FAILURE);
    }
    auto flux = [&](int tag) {
        std::vector<double> ys = {0.0, H};
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == tag) ys.push_back(mesh.y[i]);
        std::sort(ys.begin(), ys.end());
        auto prof = [H](double y) { return y * (H - y); };
        double q = 0.0;
        for (size_t k = 1; k < ys.size(); ++k) q += 0.5 * (ys[k] - ys[k - 1]) * (prof(ys[k]) + prof(ys[k - 1]));
        return q;
    };
    scale = flux(BC_INLET) / flux(BC_OUTLET);
}

// src/cfa/device.cuh:179-199: outlet values of a transferred field scaled by its own inlet over
// outlet trapezoidal flux (corners at 0), when the outlet flux is not negligible
std::vector<double> flux_balanced(const Mesh2D& mesh, const std::vector<double>& f) {
    const int n = (int)mesh.x.size();
    std::vector<double> g = f;
    auto flux = [&](int tag) {
        std::vector<std::pair<double, double>> pts = {{0.0, 0.0}, {Sim::H_CHANNEL, 0.0}};
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == tag) pts.push_back({mesh.y[i], f[i]});
        std::sort(pts.begin(), pts.end());
        double q = 0.0;
        for (size_t k = 1; k < pts.size(); ++k)
            q += 0.5 * (pts[k].first - pts[k - 1].first) * (pts[k].second + pts[k - 1].second);
        return q;
    };
    const double qi = flux(BC_INLET), qo = flux(BC_OUTLET);
    if (std::fabs(qo) > 1.0e-14)
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == BC_OUTLET) g[i] *= qi / qo;
    return g;
}

}  // namespace

// src/cfa/device.cuh:131-228 (default outlet)
void DeviceMeshBuffers::sync_from_host(const Mesh2D& mesh, const std::vector<double>& h_u,
                                       const std::vector<double>& h_v, const std::vector<double>& h_p,
                                       const std::vector<double>& h_uo, const std::vector<double>& h_vo) {
    upload_elements(mesh);
    nnz = mesh.nnz;
    d_row_ptr.upload(mesh.csr_row_ptr);
    d_col_ind.upload(mesh.csr_col_ind);
    std::vector<int> inc_ptr, inc_elem, inc_pos;
    build_incidence(mesh, inc_ptr, inc_elem, inc_pos);
    d_inc_ptr.upload(inc_ptr);
    d_inc_elem.upload(inc_elem);
    d_inc_pos.upload(inc_pos);
    upload_fields(mesh, h_u, h_v, h_p, h_uo, h_vo);
}

void DeviceMeshBuffers::sync_from_host_device_tables(const Mesh2D& mesh, const std::vector<double>& h_u,
                                                     const std::vector<double>& h_v,
                                                     const std::vector<double>& h_p,
                                                     const std::vector<double>& h_uo,
                                                     const std::vector<double>& h_vo) {
    if (mesh.num_nodes <= 0 || mesh.num_elems <= 0) {   // nothing for the device kernels to sort
        sync_from_host(mesh, h_u, h_v, h_p, h_uo, h_vo);
        return;
    }
    upload_elements(mesh);
    nnz = mesh.nnz;
    d_row_ptr.upload(mesh.csr_row_ptr);
    d_col_ind.upload(mesh.csr_col_ind);
    const int n = n_nodes, ne = n_elems;
    d_inc_ptr.reserve(n + 1);
    d_inc_elem.reserve(3 * (size_t)ne);
    d_inc_pos.reserve(9 * (size_t)ne);
    const size_t wsb = cfa_topology_ws_bytes(n, ne);
    d_topo_ws_.reserve(std::max<size_t>(wsb, 1));
    GPU_CHECK(cfa_mesh_tables(n, ne, d_elem_v, d_row_ptr, d_col_ind, d_topo_ws_, wsb, d_inc_ptr, d_inc_elem,
                              d_inc_pos, solver_stream()));
    upload_fields(mesh, h_u, h_v, h_p, h_uo, h_vo);
}

void DeviceMeshBuffers::upload_elements(const Mesh2D& mesh) {
    n_nodes = mesh.num_nodes;
    n_elems = mesh.num_elems;
    const size_t ne = n_elems;
    std::vector<int> flat(3 * ne);
    for (size_t e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) flat[3 * e + k] = mesh.elems[e].v[k];
    d_elem_v.upload(flat);
    d_bc_tag.upload(mesh.bc_tag);
    d_x.upload(mesh.x);
    d_y.upload(mesh.y);
}

bool DeviceMeshBuffers::sync_from_host_device_topology(Mesh2D& mesh, const std::vector<double>& h_u,
                                                       const std::vector<double>& h_v,
                                                       const std::vector<double>& h_p,
                                                       const std::vector<double>& h_uo,
                                                       const std::vector<double>& h_vo, MeshQuality& q,
                                                       bool& valid) {
    mesh.num_nodes = (int)mesh.x.size();
    mesh.num_elems = (int)mesh.elems.size();
    if (mesh.num_nodes <= 0 || mesh.num_elems <= 0) return false;
    upload_elements(mesh);
    const int n = n_nodes, ne = n_elems;
    const cfa_stream_t s = solver_stream();
    const size_t cap = (size_t)n + 6 * (size_t)ne;   // upper bound of the graph's entries
    d_row_ptr.reserve(n + 1);
    d_col_ind.reserve(cap);
    d_inc_ptr.reserve(n + 1);
    d_inc_elem.reserve(3 * (size_t)ne);
    d_inc_pos.reserve(9 * (size_t)ne);
    const size_t wsb = cfa_topology_ws_bytes(n, ne);
    if (wsb == 0) return false;
    d_topo_ws_.reserve(wsb);
    d_topo_out_.reserve(8);
    d_topo_nnz_.reserve(1);
    GPU_CHECK(cfa_mesh_topology(n, ne, d_elem_v, d_topo_ws_, wsb, d_inc_ptr, d_inc_elem, d_row_ptr, d_col_ind,
                                d_inc_pos, d_topo_nnz_, s));
    int nz = 0;
    d_topo_nnz_.download(&nz, 1);
    if (nz < 0) return false;   // a row longer than the device kernel takes
    GPU_CHECK(cfa_mesh_check(n, ne, d_elem_v, d_x, d_y, d_bc_tag, d_row_ptr, d_col_ind, nz, Sim::H_CHANNEL,
                             Sim::L_CHANNEL, d_topo_ws_, wsb, d_topo_out_, s));
    double st[6];
    d_topo_out_.download(st, 6);
    mesh.nnz = nz;
    mesh.csr_row_ptr.resize(n + 1);
    mesh.csr_col_ind.resize(nz);
    d_row_ptr.download(mesh.csr_row_ptr.data(), n + 1);
    d_col_ind.download(mesh.csr_col_ind.data(), nz);
    build_boundary_topology(mesh);
    q = MeshQuality();
    q.min_area = st[3];
    q.total_area = st[4];
    q.min_angle_deg = std::acos(std::max(-1.0, std::min(1.0, st[5]))) * 180.0 / Sim::PI;
    for (const Triangle& t : mesh.elems) {
        q.max_level = std::max(q.max_level, t.level);
        q.max_gen = std::max(q.max_gen, t.gen);
    }
    valid = st[0] == 0.0 && st[1] == 0.0 && st[2] == 0.0;
    if (!valid) {   // the host check names the offending elements and edges
        MeshQuality qh;
        (void)validate_mesh(mesh, qh);
    }
    nnz = nz;
    upload_fields(mesh, h_u, h_v, h_p, h_uo, h_vo);
    return true;
}

void DeviceMeshBuffers::upload_fields(const Mesh2D& mesh, const std::vector<double>& h_u,
                                      const std::vector<double>& h_v, const std::vector<double>& h_p,
                                      const std::vector<double>& h_uo, const std::vector<double>& h_vo) {
    const size_t n = n_nodes, ne = n_elems, nz = nnz;
    for (DevArray<double>* a : {&d_x, &d_y, &d_u, &d_v, &d_u_old, &d_v_old, &d_u_prev, &d_v_prev, &d_p, &d_dp, &d_vort,
                                &d_rhs_u, &d_rhs_v, &d_rhs_p, &d_mass_lumped, &d_grad_dp_x, &d_grad_dp_y, &d_vort_num,
                                &d_div_num, &d_oss_gx, &d_oss_gy, &d_inv_m, &d_phi})
        a->reserve(n);
    d_grad_rec.reserve(4 * n);
    d_elem_error.reserve(2 * ne);
    d_elem_s.reserve(ne);
    d_A_val.reserve(nz, 256);
    d_Kp_val.reserve(nz, 256);
    d_oss_val.reserve(4 * nz, 1024);
    d_diag.reserve(6, 0);
    d_ws.reserve(8 * 1024, 0);

    std::vector<int> layer;
    for (size_t e = 0; e < ne; ++e) {
        const Triangle& t = mesh.elems[e];
        if (mesh.bc_tag[t.v[0]] == BC_CYLINDER || mesh.bc_tag[t.v[1]] == BC_CYLINDER || mesh.bc_tag[t.v[2]] == BC_CYLINDER)
            layer.push_back((int)e);
    }
    n_layer = (int)layer.size();
    d_layer.upload(layer);

    if (outlet_dirichlet && n > 0) {
        outlet_pin_and_scale(mesh, p_pin, outlet_scale);
        d_u.upload(flux_balanced(mesh, h_u));
        d_u_old.upload(flux_balanced(mesh, h_uo));
    } else {
        p_pin = -1;
        outlet_scale = 1.0;
        d_u.upload(h_u);
        d_u_old.upload(h_uo);
    }
    d_v.upload(h_v);
    d_p.upload(h_p);
    d_v_old.upload(h_vo);
    dev_copy(d_u_prev, d_u_old, n);
    dev_copy(d_v_prev, d_v_old, n);
    dev_zero(d_dp, n * sizeof(double));
    dev_zero(d_vort, n * sizeof(double));
    dev_zero(d_elem_error, 2 * ne * sizeof(double));

    n_cyl_edges = (int)mesh.cylinder_edges.size();
    std::vector<int> ce(3 * (size_t)n_cyl_edges);
    for (int k = 0; k < n_cyl_edges; ++k) {
        ce[3 * k] = mesh.cylinder_edges[k].first;
        ce[3 * k + 1] = mesh.cylinder_edges[k].second;
        ce[3 * k + 2] = mesh.cylinder_edge_elem[k];
    }
    d_cyl_edges.upload(ce);
    probe_front = find_cylinder_probe(mesh, Sim::PI);
    probe_back = find_cylinder_probe(mesh, 0.0);

    view_.n_nodes = n_nodes;
    view_.n_elems = n_elems;
    view_.nnz = nnz;
    view_.elem_v = d_elem_v;
    view_.x = d_x;
    view_.y = d_y;
    view_.bc_tag = d_bc_tag;
    view_.row_ptr = d_row_ptr;
    view_.col_ind = d_col_ind;
    view_.inc_ptr = d_inc_ptr;
    view_.inc_elem = d_inc_elem;
    view_.inc_pos = d_inc_pos;
    view_.outlet_dirichlet = outlet_dirichlet && n > 0 ? 1 : 0;
    view_.p_pin = p_pin;
    view_.outlet_scale = outlet_scale;
}

void DeviceMeshBuffers::copy_solution_to_host(std::vector<double>& h_u, std::vector<double>& h_v,
                                              std::vector<double>& h_p, std::vector<double>& h_uo,
                                              std::vector<double>& h_vo, std::vector<double>& h_vort,
                                              std::vector<double>& h_elem_err) {
    const size_t n = n_nodes;
    h_u.resize(n); h_v.resize(n); h_p.resize(n); h_uo.resize(n); h_vo.resize(n); h_vort.resize(n);
    h_elem_err.resize(2 * (size_t)n_elems);
    d_u.download(h_u.data(), n);
    d_v.download(h_v.data(), n);
    d_p.download(h_p.data(), n);
    d_u_old.download(h_uo.data(), n);
    d_v_old.download(h_vo.data(), n);
    d_vort.download(h_vort.data(), n);
    d_elem_error.download(h_elem_err.data(), 2 * (size_t)n_elems);
}

// src/cfa/device.cuh:265-281
void compute_error_indicator(DeviceMeshBuffers& g, double inflow_peak) {
    GPU_CHECK(cfa_error_indicator(g.view(), g.d_u, g.d_v, inflow_peak, g.d_mass_lumped, g.d_grad_rec,
                                  g.d_elem_error, solver_stream()));
}

void apply_oss_projection_matrix_free(DeviceMeshBuffers& g, const double* in, double* out) {
    GPU_CHECK(cfa_oss_apply_matrix_free(g.view(), g.d_elem_s, g.d_mass_lumped, in, out, g.d_oss_gx, g.d_oss_gy,
                                        solver_stream()));
}

void assemble_oss_projection(DeviceMeshBuffers& g) {
    GPU_CHECK(cfa_assemble_oss_operator(g.view(), g.d_elem_s, g.d_mass_lumped, g.d_oss_val, g.d_inv_m,
                                        solver_stream()));
}

void apply_oss_projection(DeviceMeshBuffers& g, const double* in, double* out) {
    GPU_CHECK(cfa_oss_apply(g.view(), g.d_oss_val, g.d_inv_m, in, out, g.d_oss_gx, g.d_oss_gy, solver_stream()));
}

// src/cfa/device.cuh:337-408
ReprojectionAux::ReprojectionAux() {
    GPU_CHECK(cfa_stream_create(&s));
    GPU_CHECK(cfa_timer_create(&ready));
}

ReprojectionAux::~ReprojectionAux() {
    (void)cfa_stream_sync(s);
    (void)cfa_timer_destroy(ready);
    (void)cfa_stream_destroy(s);
}

int project_scheme_consistent(DeviceMeshBuffers& g, GpuLinearSolver& solver, double dt_proj, double stab,
                              AmgPreconditioner* amg, int& nonconverged, const std::function<void()>& build_amg,
                              bool both_levels, double* phi_out, double tol, ReprojectionAux* aux) {
    cfa_stream_t s = solver_stream();
    const bool oss = stab > 0.0;
    const cfa_mesh_t* m = g.view();
    GPU_CHECK(cfa_lumped_mass(m, g.d_mass_lumped, s));
    if (oss) GPU_CHECK(cfa_oss_weights(m, g.d_u, g.d_v, dt_proj, Sim::RHO, Sim::NU, stab, g.d_elem_s, s));
    const cfa_oss_op_t op = g.oss_op();
    int iters = 0;
    if (both_levels && amg && aux) {
        // The two projections do not depend on each other (the correction leaves p unchanged:
        // p_update 0): both systems are assembled, the hierarchy built from the first, and the two
        // PCG solves run at the same time, the second on its own stream, solver, right-hand side,
        // solution, OSS scratch and hierarchy work vectors; then the two corrections in order. Every
        // solve and correction sees the values it sees when they run one after the other.
        const int n = g.n_nodes;
        for (DevArray<double>* v : {&aux->rhs, &aux->div_num, &aux->dp, &aux->gx, &aux->gy}) v->reserve(n, 64);
        double* rhs[2] = {g.d_rhs_p, aux->rhs};
        double* div[2] = {g.d_div_num, aux->div_num};
        double* dp[2] = {g.d_dp, aux->dp};
        double* uv[2][2] = {{g.d_u_old, g.d_v_old}, {g.d_u, g.d_v}};
        if (oss) GPU_CHECK(cfa_scalar_gradient(m, g.d_p, g.d_elem_s, g.d_grad_dp_x, g.d_grad_dp_y, s));
        for (int f = 0; f < 2; ++f) {
            dev_zero(dp[f], n * sizeof(double));
            GPU_CHECK(cfa_assemble_pressure(m, uv[f][0], uv[f][1], g.d_p, g.d_grad_dp_x, g.d_grad_dp_y,
                                            g.d_mass_lumped, g.d_elem_s, dt_proj, Sim::RHO, stab, g.d_Kp_val, rhs[f],
                                            div[f], dp[f], oss ? g.d_oss_val.ptr : nullptr, g.d_inv_m, s));
            g.pressure_outlet_bc(g.d_mass_lumped, rhs[f], s);
        }
        if (build_amg) build_amg();
        else amg->setup(n, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
        const cfa_amg_t* view2 = amg->second_view();
        cfa_oss_op_t op2 = op;
        op2.gx = aux->gx;
        op2.gy = aux->gy;
        aux->solver.allocate(n);
        solver.allocate(n);
        GPU_CHECK(cfa_timer_record(aux->ready, s));
        GPU_CHECK(cfa_stream_wait(aux->s, aux->ready));
        int it1 = 0;
        std::thread second([&] {
            it1 = aux->solver.solve_pcg_fused(n, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, rhs[1], dp[1], 4000, tol,
                                              oss ? &op2 : nullptr, view2, aux->s);
        });
        const int it0 = solver.solve_pcg_fused(n, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, rhs[0], dp[0], 4000, tol,
                                               oss ? &op : nullptr, amg);
        second.join();
        iters = it0 + it1;
        nonconverged += (solver.converged ? 0 : 1) + (aux->solver.converged ? 0 : 1);
        double* vort[2] = {g.d_oss_gx, g.d_vort};
        for (int f = 0; f < 2; ++f)
            GPU_CHECK(cfa_correction(m, dt_proj, Sim::RHO, 0.0, 0.0, div[f], g.d_mass_lumped, dp[f], uv[f][0],
                                     uv[f][1], g.d_p, vort[f], g.d_grad_dp_x, g.d_grad_dp_y, g.d_vort_num, s));
        if (phi_out) dev_copy(phi_out, dp[1], n);
        dev_zero(g.d_dp, n * sizeof(double));
        return iters;
    }
    double* fields[2][3] = {{g.d_u_old, g.d_v_old, g.d_oss_gx}, {g.d_u, g.d_v, g.d_vort}};
    const int first = both_levels ? 0 : 1;
    for (int f = first; f < 2; ++f) {
        double* u = fields[f][0];
        double* v = fields[f][1];
        dev_zero(g.d_dp, g.n_nodes * sizeof(double));
        if (oss) GPU_CHECK(cfa_scalar_gradient(m, g.d_p, g.d_elem_s, g.d_grad_dp_x, g.d_grad_dp_y, s));
        GPU_CHECK(cfa_assemble_pressure(m, u, v, g.d_p, g.d_grad_dp_x, g.d_grad_dp_y, g.d_mass_lumped, g.d_elem_s,
                                        dt_proj, Sim::RHO, stab, g.d_Kp_val, g.d_rhs_p, g.d_div_num, g.d_dp,
                                        oss ? g.d_oss_val.ptr : nullptr, g.d_inv_m, s));
        g.pressure_outlet_bc(g.d_mass_lumped, g.d_rhs_p, s);
        if (amg && f == first) {
            if (build_amg) build_amg();
            else amg->setup(g.n_nodes, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
        }
        if (amg) {
            iters += solver.solve_pcg_fused(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000,
                                            tol, oss ? &op : nullptr, amg);
        } else {
            std::function<void(const double*, double*)> oss_fn = [&g](const double* in, double* out) {
                apply_oss_projection(g, in, out);
            };
            iters += solver.solve_pcg(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000, tol,
                                      oss ? &oss_fn : nullptr, nullptr);
        }
        nonconverged += solver.converged ? 0 : 1;
        GPU_CHECK(cfa_correction(m, dt_proj, Sim::RHO, 0.0, 0.0, g.d_div_num, g.d_mass_lumped, g.d_dp, u, v, g.d_p,
                                 fields[f][2], g.d_grad_dp_x, g.d_grad_dp_y, g.d_vort_num, s));
    }
    if (phi_out) dev_copy(phi_out, g.d_dp, g.n_nodes);
    dev_zero(g.d_dp, g.n_nodes * sizeof(double));
    return iters;
}

// src/cfa/device.cuh:418-464
int project_divergence_free(DeviceMeshBuffers& g, GpuLinearSolver& solver, double dt, AmgPreconditioner* amg,
                            bool setup_amg, double* u, double* v, double* vort_out, double tol) {
    cfa_stream_t s = solver_stream();
    const cfa_mesh_t* m = g.view();
    dev_zero(g.d_dp, g.n_nodes * sizeof(double));
    GPU_CHECK(cfa_lumped_mass(m, g.d_mass_lumped, s));
    GPU_CHECK(cfa_assemble_pressure(m, u, v, g.d_p, nullptr, nullptr, g.d_mass_lumped, nullptr, dt, Sim::RHO, 0.0,
                                    g.d_Kp_val, g.d_rhs_p, g.d_div_num, g.d_dp, nullptr, nullptr, s));
    g.pressure_outlet_bc(g.d_mass_lumped, g.d_rhs_p, s);
    int iters;
    if (amg) {
        if (setup_amg) amg->setup(g.n_nodes, g.nnz, g.d_row_ptr, g.d_col_ind, g.d_Kp_val);
        iters = solver.solve_pcg_fused(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000, tol,
                                       nullptr, amg);
    } else {
        iters = solver.solve_pcg(g.n_nodes, g.d_row_ptr, g.d_col_ind, g.d_Kp_val, g.d_rhs_p, g.d_dp, 4000, tol);
    }
    GPU_CHECK(cfa_correction(m, dt, Sim::RHO, 0.0, 0.0, g.d_div_num, g.d_mass_lumped, g.d_dp, u, v, g.d_p, vort_out,
                             g.d_grad_dp_x, g.d_grad_dp_y, g.d_vort_num, s));
    dev_zero(g.d_dp, g.n_nodes * sizeof(double));
    return iters;
}

// src/cfa/device.cuh:466-486
void set_initial_condition(const Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                           std::vector<double>& h_p, double inflow_peak) {
    h_u.assign(mesh.num_nodes, 0.0);
    h_v.assign(mesh.num_nodes, 0.0);
    h_p.assign(mesh.num_nodes, 0.0);
    for (int i = 0; i < mesh.num_nodes; ++i) {
        const double yi = mesh.y[i];
        const double prof = 4.0 * inflow_peak * yi * (Sim::H_CHANNEL - yi) / (Sim::H_CHANNEL * Sim::H_CHANNEL);
        if (mesh.bc_tag[i] == BC_INLET) {
            h_u[i] = prof;
        } else if (mesh.bc_tag[i] == BC_INTERIOR || mesh.bc_tag[i] == BC_OUTLET) {
            const double r = std::hypot(mesh.x[i] - Sim::X_CYL, yi - Sim::Y_CYL);
            const double damp = 1.0 - std::exp(-std::max(0.0, r - Sim::R_CYL) / 0.03);
            h_u[i] = damp * prof;
        }
    }
}

void print_amr_report(const char* tag, const AmrStats& st, const Mesh2D& mesh, const MeshQuality& q,
                      int proj_iters) {
    std::cout << "  --> [AMR " << tag << "] marked refine " << st.marked_refine << " / coarsen " << st.marked_coarsen
              << " | bisections " << st.bisections << " (+" << st.closure_bisections << " closure)"
              << " | removed vertices " << st.removed_vertices << "\n"
              << "      mesh: " << mesh.num_nodes << " nodes, " << mesh.num_elems << " elements"
              << " (" << st.leaves_before << " -> " << st.leaves_after << "), CSR nnz = " << mesh.nnz
              << " | max level " << q.max_level << " | min angle " << std::setprecision(2) << q.min_angle_deg
              << " deg | adapt " << std::setprecision(1) << st.elapsed_ms << " ms";
    if (proj_iters >= 0) std::cout << " | div-projection PCG " << proj_iters << " it";
    std::cout << std::endl << std::setprecision(4);
}

void RemeshProfile::report(double total_ms) const {
    if (count == 0) return;
    auto item = [&](const char* name, double ms) {
        std::cout << "    " << std::left << std::setw(26) << name << std::right << std::setw(9) << std::setprecision(2)
                  << ms / 1000.0 << " s  " << std::setw(6) << std::setprecision(1) << 100.0 * ms / total_ms << " %  "
                  << std::setw(8) << std::setprecision(1) << ms / count << " ms/remesh\n";
    };
    std::cout << "[Timing] remesh breakdown over " << count << " remeshes:\n" << std::fixed;
    if (overlapped > 0.0) {
        total_ms = indicator + overlapped + join_wait + transfer + upload + colouring + amg_setup + projection;
        std::cout << "    (--amr-overlap: adapt ... validate ran in a worker thread, " << std::setprecision(2)
                  << overlapped / 1000.0 << " s, hidden behind stepping except the join wait)\n";
    }
    item("indicators + download", indicator);
    item("adapt: look-ahead", lookahead);
    item("adapt: marking", mark);
    item("adapt: neighbours/buffer", neighbours);
    item("adapt: coarsening", coarsen);
    item("adapt: refinement", refine);
    item("extract", extract);
    item("CSR build", csr);
    item("validate_mesh", validate);
    item("upload", upload);
    if (device_topology > 0.0) item("upload + CSR/checks on GPU", device_topology);
    item("ILU colouring", colouring);
    item("re-projection AMG setup", amg_setup);
    item("re-projection solve", projection);
    if (overlapped > 0.0) {
        item("overlap: join wait", join_wait);
        item("overlap: transfer", transfer);
    }
}

}  // namespace cfa
