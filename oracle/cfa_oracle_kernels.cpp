// ORACLE (test infrastructure only): element kernels of src/cfa/fem_kernels.cuh restated as
// serial element loops. Each function accumulates into caller-zeroed outputs, as the reference
// kernels do; contributions are added in element order (the reference adds them with atomics
// in an unspecified order, so agreement is to round-off).
#include <algorithm>
#include <cmath>

#include "cfa_oracle.hpp"

namespace oracle {

using cfa::BC_CYLINDER;
using cfa::BC_INLET;
using cfa::BC_INTERIOR;
using cfa::BC_OUTLET;
using cfa::BC_WALL;

namespace {
struct Geo {
    double area, dNdx[3], dNdy[3];
};
// P1 gradients of an element (src/cfa/fem_kernels.cuh:81-87); false for a degenerate element
inline bool element_geometry(const int* n, const double* x, const double* y, Geo& g) {
    const double x0 = x[n[0]], y0 = y[n[0]], x1 = x[n[1]], y1 = y[n[1]], x2 = x[n[2]], y2 = y[n[2]];
    const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    g.area = 0.5 * std::fabs(detJ);
    if (g.area < 1.0e-14) return false;
    const double inv2A = 1.0 / (2.0 * g.area);
    g.dNdx[0] = (y1 - y2) * inv2A;  g.dNdx[1] = (y2 - y0) * inv2A;  g.dNdx[2] = (y0 - y1) * inv2A;
    g.dNdy[0] = (x2 - x1) * inv2A;  g.dNdy[1] = (x0 - x2) * inv2A;  g.dNdy[2] = (x1 - x0) * inv2A;
    return true;
}
}  // namespace

// src/cfa/fem_kernels.cuh:12-28
int find_csr(const int* ci, int s, int e, int col) {
    int lo = s, hi = e - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int c = ci[mid];
        if (c == col) return mid;
        if (c < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

// src/cfa/fem_kernels.cuh:299-346 (closed form with the sympy CSE constants of the reference)
MmsValues mms_exact(double x, double y, double t) {
    MmsValues r;
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
    const double c24 = std::cos(c12);
    const double c25 = 9.8696044010893586;
    const double c26 = c14 * c25;
    const double c27 = 0.00297441998810232 * c26 * (c8 - 2.0 * c9);
    const double c28 = 1.0 / c25;
    const double c29 = std::pow(c14, 2);
    const double c30 = 5.9488399762046401 * c25;
    const double c31 = c5 * (-2.97441998810232 * c25 * c9 - c30 * c8);
    const double c32 = 0.1681 * c28 * c29;
    const double c33 = c16 * (-c2 * c30 + 5.9488399762046401 * c25 * c4);
    const double c34 = c10 * c15 * c19;
    r.u = c11 * c15;
    r.v = -c15 * c20;
    r.p = c23 * c9;
    r.fx = 0.1681 * std::pow(c10, 2) * c19 * c28 * c29 * c5 + 2.05 * c10 * c24 * c5 - c20 * c31 * c32 -
           c21 * c6 * c9 * std::cos(c22) - c27 * c5 - c27 * (c2 - c4);
    r.fy = c10 * c16 * std::pow(c19, 2) * c32 - c11 * c32 * c33 + 0.0059488399762046401 * c16 * c26 * (c17 - 2.0 * c18) -
           2.05 * c20 * c24 - c23 * c6 * c8 + 0.00148720999405116 * c26 * (2.0 * c17 - c18) * (2.0 * c8 + c9);
    r.ux = c34;
    r.uy = c15 * c31;
    r.vx = -c15 * c33;
    r.vy = -c34;
    return r;
}

// src/cfa/fem_kernels.cuh:40-181
void assemble_momentum(int ne, const int* ev, const double* x, const double* y,
                       const double* u_old, const double* v_old, const double* u_prev,
                       const double* v_prev, const double* p_old, double nu, double rho,
                       double a0, double a1, double a2, double c1, double c2, double tau_time,
                       double skew, const int* rp, const int* ci, double* A, double* rhs_u,
                       double* rhs_v, double* mass, const double* fx, const double* fy) {
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        const double area = g.area;
        double uh[3], vh[3], wu_bar = 0.0, wv_bar = 0.0, div_w = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double un = u_old[n[k]], vn = v_old[n[k]], um = u_prev[n[k]], vm = v_prev[n[k]];
            uh[k] = a1 * un + a2 * um;
            vh[k] = a1 * vn + a2 * vm;
            const double wu = c1 * un + c2 * um, wv = c1 * vn + c2 * vm;
            wu_bar += wu / 3.0;
            wv_bar += wv / 3.0;
            div_w += wu * g.dNdx[k] + wv * g.dNdy[k];
        }
        const double vel = std::sqrt(wu_bar * wu_bar + wv_bar * wv_bar) + 1.0e-12;
        const double hK = std::sqrt(2.0 * area);
        const double tt = 2.0 / tau_time, tc = 2.0 * vel / hK, td = 4.0 * nu / (hK * hK);
        const double tau = 1.0 / std::sqrt(tt * tt + tc * tc + 9.0 * td * td);
        const double pe[3] = {p_old[n[0]], p_old[n[1]], p_old[n[2]]};
        const double dpdx = (pe[0] * g.dNdx[0] + pe[1] * g.dNdx[1] + pe[2] * g.dNdx[2]) / rho;
        const double dpdy = (pe[0] * g.dNdy[0] + pe[1] * g.dNdy[1] + pe[2] * g.dNdy[2]) / rho;
        double ug[3];
        for (int i = 0; i < 3; ++i) ug[i] = wu_bar * g.dNdx[i] + wv_bar * g.dNdy[i];
        const double m_lump = area / 3.0;
        for (int i = 0; i < 3; ++i) {
            const int row = n[i];
            mass[row] += m_lump;
            double f_u = 0.0, f_v = 0.0;
            for (int j = 0; j < 3; ++j) {
                const int col = n[j];
                const double M = (i == j) ? (area / 6.0) : (area / 12.0);
                const double K = area * (g.dNdx[i] * g.dNdx[j] + g.dNdy[i] * g.dNdy[j]);
                const double C = (area / 3.0) * ug[j] + skew * div_w * M;
                const double S = area * tau * ug[i] * ug[j];
                const double Ms = (area / 3.0) * tau * ug[i];
                const double Aij = a0 * (M + Ms) + nu * K + C + S;
                const int idx = find_csr(ci, rp[row], rp[row + 1], col);
                if (idx >= 0) A[idx] += Aij;
                f_u += (M + Ms) * uh[j];
                f_v += (M + Ms) * vh[j];
                if (fx) {
                    f_u += (M + Ms) * fx[col];
                    f_v += (M + Ms) * fy[col];
                }
            }
            const double tw = (area / 3.0) + area * tau * ug[i];
            f_u -= tw * dpdx;
            f_v -= tw * dpdy;
            rhs_u[row] += f_u;
            rhs_v[row] += f_v;
        }
    }
}

// src/cfa/fem_kernels.cuh:411-453 (default outlet: p = 0, no velocity condition there)
void apply_bc_momentum(int n, const int* tag, const double* x, const double* y, double u_max,
                       double mms_time, const int* rp, const int* ci, double* A, double* rhs_u,
                       double* rhs_v, double* u, double* v, const Outlet& outlet) {
    const double H = cfa::Sim::H_CHANNEL;
    for (int i = 0; i < n; ++i) {
        const int t = tag[i];
        const bool outlet_row = t == BC_OUTLET && outlet.dirichlet;
        if (!(t == BC_INLET || t == BC_WALL || t == BC_CYLINDER || outlet_row)) continue;
        double ub = 0.0, vb = 0.0;
        if (mms_time >= 0.0) {
            const MmsValues r = mms_exact(x[i], y[i], mms_time);
            ub = r.u;
            vb = r.v;
        } else if (t == BC_INLET) {
            const double yi = y[i];
            ub = 4.0 * u_max * yi * (H - yi) / (H * H);
        } else if (outlet_row) {   // :440, the parabola times the flux factor
            const double yi = y[i];
            ub = 4.0 * u_max * yi * (H - yi) / (H * H) * outlet.scale;
        }
        for (int k = rp[i]; k < rp[i + 1]; ++k) A[k] = (ci[k] == i) ? 1.0 : 0.0;
        rhs_u[i] = ub;  rhs_v[i] = vb;
        u[i] = ub;      v[i] = vb;
    }
}

// src/cfa/fem_kernels.cuh:348-355
void mms_forcing(int n, const double* x, const double* y, double t, double* fx, double* fy) {
    for (int i = 0; i < n; ++i) {
        const MmsValues r = mms_exact(x[i], y[i], t);
        fx[i] = r.fx;
        fy[i] = r.fy;
    }
}

// src/cfa/fem_kernels.cuh:697-719
void compute_oss_weights(int ne, const int* ev, const double* x, const double* y, const double* us,
                         const double* vs, double dt, double rho, double nu, double stab, double* s) {
    for (int e = 0; e < ne; ++e) {
        const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
        const double area = 0.5 * std::fabs((x[n1] - x[n0]) * (y[n2] - y[n0]) - (x[n2] - x[n0]) * (y[n1] - y[n0]));
        const double h = std::sqrt(2.0 * area);
        const double ub = (us[n0] + us[n1] + us[n2]) / 3.0;
        const double vb = (vs[n0] + vs[n1] + vs[n2]) / 3.0;
        const double tau = stab / (rho * (4.0 * nu / (h * h) + 2.0 * std::sqrt(ub * ub + vb * vb) / h) + 1.0e-300);
        s[e] = rho * tau / dt;
    }
}

// src/cfa/fem_kernels.cuh:726-757 (s == nullptr: unit weights)
void accumulate_scalar_gradient(int ne, const int* ev, const double* x, const double* y,
                                const double* f, const double* s, double* gx, double* gy) {
    for (int e = 0; e < ne; ++e) {
        const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
        const double x0 = x[n0], y0 = y[n0], x1 = x[n1], y1 = y[n1], x2 = x[n2], y2 = y[n2];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double area = 0.5 * std::fabs(detJ);
        if (area < 1.0e-14) continue;
        const double inv2A = 1.0 / (2.0 * area);
        const double g_x = (f[n0] * (y1 - y2) + f[n1] * (y2 - y0) + f[n2] * (y0 - y1)) * inv2A;
        const double g_y = (f[n0] * (x2 - x1) + f[n1] * (x0 - x2) + f[n2] * (x1 - x0)) * inv2A;
        const double w = (s ? std::sqrt(s[e]) : 1.0) * area / 3.0;
        gx[n0] += w * g_x;  gy[n0] += w * g_y;
        gx[n1] += w * g_x;  gy[n1] += w * g_y;
        gx[n2] += w * g_x;  gy[n2] += w * g_y;
    }
}

// src/cfa/fem_kernels.cuh:201-282
void assemble_pressure_poisson(int ne, const int* ev, const double* x, const double* y,
                               const double* us, const double* vs, const double* p,
                               const double* gpx, const double* gpy, const double* mass,
                               const double* s_el, double dt, double rho, double stab, const int* rp,
                               const int* ci, double* K, double* rhs, double* div_num) {
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        const double area = g.area;
        const double div_u = us[n[0]] * g.dNdx[0] + us[n[1]] * g.dNdx[1] + us[n[2]] * g.dNdx[2] +
                             vs[n[0]] * g.dNdy[0] + vs[n[1]] * g.dNdy[1] + vs[n[2]] * g.dNdy[2];
        const double rc = -(rho / dt) * (area / 3.0) * div_u;
        double s = 0.0, sx = 0.0, sy = 0.0, gx = 0.0, gy = 0.0;
        if (stab > 0.0) {
            s = s_el[e];
            const double rs = std::sqrt(s);
            gx = p[n[0]] * g.dNdx[0] + p[n[1]] * g.dNdx[1] + p[n[2]] * g.dNdx[2];
            gy = p[n[0]] * g.dNdy[0] + p[n[1]] * g.dNdy[1] + p[n[2]] * g.dNdy[2];
            for (int k = 0; k < 3; ++k) {
                const double inv_m = 1.0 / std::fmax(mass[n[k]], 1.0e-14);
                sx += gpx[n[k]] * inv_m;
                sy += gpy[n[k]] * inv_m;
            }
            sx *= rs / 3.0;
            sy *= rs / 3.0;
        }
        for (int i = 0; i < 3; ++i) {
            const int row = n[i];
            for (int j = 0; j < 3; ++j) {
                const double Kij = area * (g.dNdx[i] * g.dNdx[j] + g.dNdy[i] * g.dNdy[j]);
                const int idx = find_csr(ci, rp[row], rp[row + 1], n[j]);
                if (idx >= 0) K[idx] += (1.0 + s) * Kij;
            }
            const double stab_i = s * area * (g.dNdx[i] * gx + g.dNdy[i] * gy) -
                                  area * (g.dNdx[i] * sx + g.dNdy[i] * sy);
            rhs[row] += rc - stab_i;
            div_num[row] += (area / 3.0) * div_u;
        }
    }
}

// src/cfa/fem_kernels.cuh:460-490 (Dirichlet rows: outlet nodes, or the pinned node)
void apply_bc_pressure(int n, const int* tag, const int* rp, const int* ci, double* K, double* rhs,
                       double* dp, const Outlet& outlet) {
    for (int i = 0; i < n; ++i) {
        if (outlet.p_dirichlet(i, tag[i])) {
            for (int k = rp[i]; k < rp[i + 1]; ++k) K[k] = (ci[k] == i) ? 1.0 : 0.0;
            rhs[i] = 0.0;
            dp[i] = 0.0;
        } else {
            for (int k = rp[i]; k < rp[i + 1]; ++k)
                if (ci[k] != i && outlet.p_dirichlet(ci[k], tag[ci[k]])) K[k] = 0.0;
        }
    }
}

// src/cfa/fem_kernels.cuh:493-510: k_sum reduces blocks of 256 threads by the tree
// s[t] += s[t + w], w = 128 ... 1, and adds each block's s[0] to a zeroed total (ascending
// block order here); k_subtract_mean then forms f_i -= (sum f / sum m) m_i
void subtract_mean(int n, double* f, const double* m) {
    auto total = [n](const double* a) {
        double acc = 0.0;
        double s[256];
        for (int b0 = 0; b0 < n; b0 += 256) {
            for (int t = 0; t < 256; ++t) s[t] = b0 + t < n ? a[b0 + t] : 0.0;
            for (int w = 128; w > 0; w /= 2)
                for (int t = 0; t < w; ++t) s[t] += s[t + w];
            acc += s[0];
        }
        return acc;
    };
    const double sf = total(f), sm = total(m);
    for (int i = 0; i < n; ++i) f[i] -= sf / sm * m[i];
}

// src/cfa/device.cuh:153-172
Outlet outlet_dirichlet_setup(const Mesh2D& mesh) {
    const double H = cfa::Sim::H_CHANNEL;
    const int n = mesh.num_nodes;
    Outlet o;
    o.dirichlet = true;
    for (int i = 0; i < n; ++i) {
        if (mesh.bc_tag[i] != BC_OUTLET) continue;
        if (o.pin < 0 || std::fabs(mesh.y[i] - 0.5 * H) < std::fabs(mesh.y[o.pin] - 0.5 * H)) o.pin = i;
    }
    auto unit_flux = [&](int t) {
        std::vector<double> ys{0.0, H};
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == t) ys.push_back(mesh.y[i]);
        std::sort(ys.begin(), ys.end());
        double q = 0.0;
        for (size_t k = 1; k < ys.size(); ++k)
            q += 0.5 * (ys[k] - ys[k - 1]) * (ys[k] * (H - ys[k]) + ys[k - 1] * (H - ys[k - 1]));
        return q;
    };
    o.scale = unit_flux(BC_INLET) / unit_flux(BC_OUTLET);
    return o;
}

// src/cfa/device.cuh:179-199
Vec flux_balanced(const Mesh2D& mesh, const Vec& f) {
    const int n = mesh.num_nodes;
    auto flux = [&](int t) {
        std::vector<std::pair<double, double>> pts{{0.0, 0.0}, {cfa::Sim::H_CHANNEL, 0.0}};
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == t) pts.emplace_back(mesh.y[i], f[i]);
        std::sort(pts.begin(), pts.end());
        double q = 0.0;
        for (size_t k = 1; k < pts.size(); ++k)
            q += 0.5 * (pts[k].first - pts[k - 1].first) * (pts[k].second + pts[k - 1].second);
        return q;
    };
    const double qi = flux(BC_INLET), qo = flux(BC_OUTLET);
    Vec g = f;
    if (std::fabs(qo) > 1.0e-14)
        for (int i = 0; i < n; ++i)
            if (mesh.bc_tag[i] == BC_OUTLET) g[i] *= qi / qo;
    return g;
}

// src/cfa/fem_kernels.cuh:890-909
void assemble_lumped_mass(int ne, const int* ev, const double* x, const double* y, double* m) {
    for (int e = 0; e < ne; ++e) {
        const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
        const double detJ = (x[n1] - x[n0]) * (y[n2] - y[n0]) - (x[n2] - x[n0]) * (y[n1] - y[n0]);
        const double w = std::fabs(detJ) / 6.0;
        m[n0] += w;  m[n1] += w;  m[n2] += w;
    }
}

// src/cfa/fem_kernels.cuh:807-847: oss_val = [Bx | By | BTx | BTy]
void assemble_oss_operator(int ne, const int* ev, const double* x, const double* y, const double* s,
                           const int* rp, const int* ci, int nnz, double* oss_val) {
    double* Bx = oss_val;
    double* By = oss_val + nnz;
    double* BTx = oss_val + 2 * (size_t)nnz;
    double* BTy = oss_val + 3 * (size_t)nnz;
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        const double w = std::sqrt(s[e]) * g.area / 3.0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                const int k = n[a], j = n[b];
                const int kj = find_csr(ci, rp[k], rp[k + 1], j);
                const int jk = find_csr(ci, rp[j], rp[j + 1], k);
                if (kj < 0 || jk < 0) continue;
                Bx[kj] += w * g.dNdx[b];
                By[kj] += w * g.dNdy[b];
                BTx[jk] += w * g.dNdx[b];
                BTy[jk] += w * g.dNdy[b];
            }
    }
}

// src/cfa/fem_kernels.cuh:855-885: out -= B^T M^-1 B in, Dirichlet (outlet) rows untouched
void oss_apply(int n, const int* rp, const int* ci, int nnz, const double* oss_val,
               const double* inv_m, const int* tag, const double* in, double* out, const Outlet& outlet) {
    std::vector<double> gx(n), gy(n);
    for (int k = 0; k < n; ++k) {
        double sx = 0.0, sy = 0.0;
        for (int a = rp[k]; a < rp[k + 1]; ++a) {
            const double xj = in[ci[a]];
            sx += oss_val[a] * xj;
            sy += oss_val[nnz + a] * xj;
        }
        gx[k] = sx * inv_m[k];
        gy[k] = sy * inv_m[k];
    }
    const double* BTx = oss_val + 2 * (size_t)nnz;
    const double* BTy = oss_val + 3 * (size_t)nnz;
    for (int i = 0; i < n; ++i) {
        if (outlet.p_dirichlet(i, tag[i])) continue;
        double s = 0.0;
        for (int a = rp[i]; a < rp[i + 1]; ++a) s += BTx[a] * gx[ci[a]] + BTy[a] * gy[ci[a]];
        out[i] -= s;
    }
}

// src/cfa/fem_kernels.cuh:763-797 with device.cuh:287-297: the matrix-free element form
void oss_apply_matrix_free(int ne, int n, const int* ev, const double* x, const double* y,
                           const int* tag, const double* s, const double* mass, const double* in,
                           double* out, const Outlet& outlet) {
    std::vector<double> gx(n, 0.0), gy(n, 0.0);
    accumulate_scalar_gradient(ne, ev, x, y, in, s, gx.data(), gy.data());
    for (int e = 0; e < ne; ++e) {
        const int* nd = ev + 3 * e;
        Geo g;
        if (!element_geometry(nd, x, y, g)) continue;
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double inv_m = 1.0 / std::fmax(mass[nd[k]], 1.0e-14);
            sx += gx[nd[k]] * inv_m;
            sy += gy[nd[k]] * inv_m;
        }
        const double w = std::sqrt(s[e]) * g.area / 3.0;
        for (int i = 0; i < 3; ++i) {
            if (outlet.p_dirichlet(nd[i], tag[nd[i]])) continue;
            out[nd[i]] += -w * (g.dNdx[i] * sx + g.dNdy[i] * sy);
        }
    }
}

// src/cfa/fem_kernels.cuh:536-582
void accumulate_grad_dp_and_vorticity(int ne, const int* ev, const double* x, const double* y,
                                      const double* dp, const double* u, const double* v,
                                      double* gx, double* gy, double* vort_num) {
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        const double dpdx = dp[n[0]] * g.dNdx[0] + dp[n[1]] * g.dNdx[1] + dp[n[2]] * g.dNdx[2];
        const double dpdy = dp[n[0]] * g.dNdy[0] + dp[n[1]] * g.dNdy[1] + dp[n[2]] * g.dNdy[2];
        const double dvdx = v[n[0]] * g.dNdx[0] + v[n[1]] * g.dNdx[1] + v[n[2]] * g.dNdx[2];
        const double dudy = u[n[0]] * g.dNdy[0] + u[n[1]] * g.dNdy[1] + u[n[2]] * g.dNdy[2];
        const double om = dvdx - dudy;
        const double w = g.area / 3.0;
        for (int i = 0; i < 3; ++i) {
            gx[n[i]] += w * dpdx;
            gy[n[i]] += w * dpdy;
            vort_num[n[i]] += w * om;
        }
    }
}

// src/cfa/fem_kernels.cuh:592-622 (outlet velocities are corrected with the default outlet only)
void velocity_pressure_correction(int n, const int* tag, double dt, double rho, double p_update,
                                  double rot, const double* div_num, const double* mass,
                                  const double* gx, const double* gy, const double* vort_num,
                                  const double* dp, double* u, double* v, double* p, double* vort,
                                  const Outlet& outlet) {
    for (int i = 0; i < n; ++i) {
        const double m = std::fmax(mass[i], 1.0e-14);
        vort[i] = vort_num[i] / m;
        p[i] += p_update * (dp[i] - rot * div_num[i] / m);
        if (tag[i] == BC_INTERIOR || (tag[i] == BC_OUTLET && !outlet.dirichlet)) {
            u[i] -= (dt / rho) * (gx[i] / m);
            v[i] -= (dt / rho) * (gy[i] / m);
        }
    }
}

// src/cfa/fem_kernels.cuh:627-649
void stiffness_apply(int ne, const int* ev, const double* x, const double* y, const double* f,
                     double c, double* out) {
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        const double gx = f[n[0]] * g.dNdx[0] + f[n[1]] * g.dNdx[1] + f[n[2]] * g.dNdx[2];
        const double gy = f[n[0]] * g.dNdy[0] + f[n[1]] * g.dNdy[1] + f[n[2]] * g.dNdy[2];
        for (int i = 0; i < 3; ++i) out[n[i]] += c * g.area * (g.dNdx[i] * gx + g.dNdy[i] * gy);
    }
}

// src/cfa/fem_kernels.cuh:1046-1123
void volume_forces(int ne, const int* ev, const int* tag, const double* x, const double* y,
                   const double* u, const double* v, const double* u_old, const double* v_old,
                   const double* u_prev, const double* v_prev, const double* p, double b0,
                   double b1, double b2, double nu, double rho, double* out) {
    for (int e = 0; e < ne; ++e) {
        const int* n = ev + 3 * e;
        double phi[3];
        bool touches = false;
        for (int i = 0; i < 3; ++i) {
            phi[i] = (tag[n[i]] == BC_CYLINDER) ? 1.0 : 0.0;
            touches = touches || phi[i] > 0.0;
        }
        if (!touches) continue;
        Geo g;
        if (!element_geometry(n, x, y, g)) continue;
        double ue[3], ve[3], ute[3], vte[3];
        double dudx = 0, dudy = 0, dvdx = 0, dvdy = 0, dphidx = 0, dphidy = 0, pbar = 0;
        for (int i = 0; i < 3; ++i) {
            ue[i] = u[n[i]];
            ve[i] = v[n[i]];
            ute[i] = b0 * ue[i] + b1 * u_old[n[i]] + b2 * u_prev[n[i]];
            vte[i] = b0 * ve[i] + b1 * v_old[n[i]] + b2 * v_prev[n[i]];
            dudx += ue[i] * g.dNdx[i];  dudy += ue[i] * g.dNdy[i];
            dvdx += ve[i] * g.dNdx[i];  dvdy += ve[i] * g.dNdy[i];
            dphidx += phi[i] * g.dNdx[i];
            dphidy += phi[i] * g.dNdy[i];
            pbar += p[n[i]] / 3.0;
        }
        double fx = 0.0, fy = 0.0;
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3;
            const double pm = 0.5 * (phi[i] + phi[j]);
            if (pm == 0.0) continue;
            const double um = 0.5 * (ue[i] + ue[j]), vm = 0.5 * (ve[i] + ve[j]);
            const double cx = um * dudx + vm * dudy, cy = um * dvdx + vm * dvdy;
            fx += (0.5 * (ute[i] + ute[j]) + cx) * pm;
            fy += (0.5 * (vte[i] + vte[j]) + cy) * pm;
        }
        fx *= g.area / 3.0;
        fy *= g.area / 3.0;
        fx += g.area * (nu * (dudx * dphidx + dudy * dphidy) - (pbar / rho) * dphidx);
        fy += g.area * (nu * (dvdx * dphidx + dvdy * dphidy) - (pbar / rho) * dphidy);
        out[0] += fx;
        out[1] += fy;
    }
}

// src/cfa/fem_kernels.cuh:657-692
void surface_forces(int nedges, const int* edges, const int* ev, const double* x, const double* y,
                    const double* u, const double* v, const double* p, double nu, double rho,
                    double xc, double yc, double* out) {
    for (int k = 0; k < nedges; ++k) {
        const int a = edges[3 * k], b = edges[3 * k + 1], e = edges[3 * k + 2];
        const int* n = ev + 3 * e;
        const double x0 = x[n[0]], y0 = y[n[0]], x1 = x[n[1]], y1 = y[n[1]], x2 = x[n[2]], y2 = y[n[2]];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double dNdx[3] = {(y1 - y2) / detJ, (y2 - y0) / detJ, (y0 - y1) / detJ};
        const double dNdy[3] = {(x2 - x1) / detJ, (x0 - x2) / detJ, (x1 - x0) / detJ};
        double ux = 0, uy = 0, vx = 0, vy = 0;
        for (int i = 0; i < 3; ++i) {
            ux += u[n[i]] * dNdx[i];  uy += u[n[i]] * dNdy[i];
            vx += v[n[i]] * dNdx[i];  vy += v[n[i]] * dNdy[i];
        }
        const double dx = x[b] - x[a], dy = y[b] - y[a];
        const double len = std::sqrt(dx * dx + dy * dy);
        double ex = dy / len, ey = -dx / len;
        if (ex * (0.5 * (x[a] + x[b]) - xc) + ey * (0.5 * (y[a] + y[b]) - yc) < 0.0) { ex = -ex; ey = -ey; }
        const double pm = 0.5 * (p[a] + p[b]) / rho;
        out[0] += len * (nu * (ux * ex + uy * ey) - pm * ex);
        out[1] += len * (nu * (vx * ex + vy * ey) - pm * ey);
    }
}

// src/cfa/fem_kernels.cuh:362-405 (7-point degree-5 rule)
void mms_errors(int ne, const int* ev, const double* x, const double* y, const double* uh,
                const double* vh, const double* ph, double t, double* out) {
    const double a1 = 0.059715871789770, b1 = 0.470142064105115, w1 = 0.132394152788506;
    const double a2 = 0.797426985353087, b2 = 0.101286507323456, w2 = 0.125939180544827;
    const double Lq[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3}, {a1, b1, b1}, {b1, a1, b1}, {b1, b1, a1},
                             {a2, b2, b2}, {b2, a2, b2}, {b2, b2, a2}};
    const double W[7] = {0.225, w1, w1, w1, w2, w2, w2};
    for (int e = 0; e < ne; ++e) {
        const int n[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        const double x0 = x[n[0]], y0 = y[n[0]], x1 = x[n[1]], y1 = y[n[1]], x2 = x[n[2]], y2 = y[n[2]];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double area = 0.5 * std::fabs(detJ);
        const double inv = 1.0 / detJ;
        const double dNdx[3] = {(y1 - y2) * inv, (y2 - y0) * inv, (y0 - y1) * inv};
        const double dNdy[3] = {(x2 - x1) * inv, (x0 - x2) * inv, (x1 - x0) * inv};
        double uhx = 0, uhy = 0, vhx = 0, vhy = 0;
        for (int k = 0; k < 3; ++k) {
            uhx += uh[n[k]] * dNdx[k];  uhy += uh[n[k]] * dNdy[k];
            vhx += vh[n[k]] * dNdx[k];  vhy += vh[n[k]] * dNdy[k];
        }
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 7; ++q) {
            const double xq = Lq[q][0] * x0 + Lq[q][1] * x1 + Lq[q][2] * x2;
            const double yq = Lq[q][0] * y0 + Lq[q][1] * y1 + Lq[q][2] * y2;
            double uq = 0, vq = 0, pq = 0;
            for (int k = 0; k < 3; ++k) {
                uq += Lq[q][k] * uh[n[k]];  vq += Lq[q][k] * vh[n[k]];  pq += Lq[q][k] * ph[n[k]];
            }
            const MmsValues r = mms_exact(xq, yq, t);
            const double w = W[q] * area;
            acc[0] += w * ((uq - r.u) * (uq - r.u) + (vq - r.v) * (vq - r.v));
            acc[1] += w * ((uhx - r.ux) * (uhx - r.ux) + (uhy - r.uy) * (uhy - r.uy) +
                           (vhx - r.vx) * (vhx - r.vx) + (vhy - r.vy) * (vhy - r.vy));
            acc[2] += w * (pq - r.p) * (pq - r.p);
            acc[3] += w * (r.u * r.u + r.v * r.v);
            acc[4] += w * (r.ux * r.ux + r.uy * r.uy + r.vx * r.vx + r.vy * r.vy);
            acc[5] += w * r.p * r.p;
        }
        for (int k = 0; k < 6; ++k) out[k] += acc[k];
    }
}

// src/cfa/device.cuh:265-281 with src/cfa/fem_kernels.cuh:921-1035: lumped mass, ZZ-recovered
// disturbance gradient, and the two indicators [eta_v | eta_i]
void error_indicator(int ne, int n, const int* ev, const double* x, const double* y,
                     const double* u, const double* v, double u_max, double* mass,
                     double* grad_rec, double* elem_error) {
    const double H = cfa::Sim::H_CHANNEL;
    std::fill(mass, mass + n, 0.0);
    std::fill(grad_rec, grad_rec + 4 * (size_t)n, 0.0);
    assemble_lumped_mass(ne, ev, x, y, mass);
    for (int e = 0; e < ne; ++e) {
        const int* nd = ev + 3 * e;
        Geo g;
        if (!element_geometry(nd, x, y, g)) continue;
        double ud[3];
        for (int i = 0; i < 3; ++i) {
            const double yi = y[nd[i]];
            ud[i] = u[nd[i]] - 4.0 * u_max * yi * (H - yi) / (H * H);
        }
        const double gr[4] = {ud[0] * g.dNdx[0] + ud[1] * g.dNdx[1] + ud[2] * g.dNdx[2],
                              ud[0] * g.dNdy[0] + ud[1] * g.dNdy[1] + ud[2] * g.dNdy[2],
                              v[nd[0]] * g.dNdx[0] + v[nd[1]] * g.dNdx[1] + v[nd[2]] * g.dNdx[2],
                              v[nd[0]] * g.dNdy[0] + v[nd[1]] * g.dNdy[1] + v[nd[2]] * g.dNdy[2]};
        const double w = g.area / 3.0;
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 4; ++c) grad_rec[(size_t)c * n + nd[i]] += w * gr[c];
    }
    for (int e = 0; e < ne; ++e) {
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        double G[3][4];
        for (int i = 0; i < 3; ++i) {
            const double inv_m = 1.0 / std::fmax(mass[nd[i]], 1.0e-14);
            for (int c = 0; c < 4; ++c) G[i][c] = grad_rec[(size_t)c * n + nd[i]] * inv_m;
        }
        const double x0 = x[nd[0]], y0 = y[nd[0]];
        const double area = 0.5 * std::fabs((x[nd[1]] - x0) * (y[nd[2]] - y0) - (x[nd[2]] - x0) * (y[nd[1]] - y0));
        double omega = 0.0, eta_i = 0.0;
        for (int i = 0; i < 3; ++i) {
            omega = std::fmax(omega, std::fabs(G[i][2] - G[i][1]));
            const int j = (i + 1) % 3;
            const double ex = x[nd[j]] - x[nd[i]], ey = y[nd[j]] - y[nd[i]];
            const double eu = (G[j][0] - G[i][0]) * ex + (G[j][1] - G[i][1]) * ey;
            const double evv = (G[j][2] - G[i][2]) * ex + (G[j][3] - G[i][3]) * ey;
            eta_i = std::fmax(eta_i, 0.125 * std::sqrt(eu * eu + evv * evv));
        }
        elem_error[e] = std::sqrt(2.0 * area) * omega;
        elem_error[ne + e] = eta_i;
    }
}

}  // namespace oracle
