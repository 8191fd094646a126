// Finite-element kernels for gfx950 (assembly, boundary rows, OSS operator, correction, forces,
// indicators, manufactured solution). Reference: src/cfa/fem_kernels.cuh.
//
// Design: the reference scatters element contributions with FP64 atomicAdd (one thread per
// element). Here every matrix/vector row is owned by one thread that gathers the contributions
// of its incident elements (node -> element table of cfa_mesh_t), in ascending element order.
// There are no atomics, the result is deterministic, and with contraction disabled it is the
// same sum, in the same order, as a serial element loop. Element quantities are recomputed by
// each of the three owners; the kernels are bound by the gathers, not by that arithmetic.
#include <mutex>

#include "cfa_device.hpp"

using namespace cfa_dev;

#pragma clang fp contract(off)

namespace {

struct Geo {
    double area, dNdx[3], dNdy[3];
};

// P1 gradients (src/cfa/fem_kernels.cuh:81-87); false for a degenerate element (|K| < 1e-14)
__device__ __forceinline__ bool geometry(const double* __restrict__ x, const double* __restrict__ y,
                                         int n0, int n1, int n2, Geo& g) {
    const double x0 = x[n0], y0 = y[n0], x1 = x[n1], y1 = y[n1], x2 = x[n2], y2 = y[n2];
    const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    g.area = 0.5 * fabs(detJ);
    if (g.area < 1.0e-14) return false;
    const double inv2A = 1.0 / (2.0 * g.area);
    g.dNdx[0] = (y1 - y2) * inv2A;  g.dNdx[1] = (y2 - y0) * inv2A;  g.dNdx[2] = (y0 - y1) * inv2A;
    g.dNdy[0] = (x2 - x1) * inv2A;  g.dNdy[1] = (x0 - x2) * inv2A;  g.dNdy[2] = (x1 - x0) * inv2A;
    return true;
}

struct Mms { double u, v, p, fx, fy, ux, uy, vx, vy; };

// Manufactured solution (src/cfa/fem_kernels.cuh:287-346): closed form with the reference's
// sympy common-subexpression constants.
__device__ Mms mms_exact(double x, double y, double t) {
    Mms r;
    const double c0 = 2.4390243902439024 * x - 5.3658536585365854;
    const double c1 = 3.1415926535897932 * c0;
    const double c2 = cos(c1);
    const double c3 = 15.324842212633138 * x - 33.714652867792903;
    const double c4 = cos(c3);
    const double c5 = c2 - 0.25 * c4;
    const double c6 = 7.6624211063165689;
    const double c7 = c6 * y;
    const double c8 = sin(c7);
    const double c9 = cos(c7);
    const double c10 = -3.8312105531582844 * c8 + 7.6624211063165689 * c9;
    const double c11 = c10 * c5;
    const double c12 = 31.415926535897932 * t;
    const double c13 = sin(c12);
    const double c14 = 0.5 * c13 + 1.0;
    const double c15 = 0.13050705333535418 * c14;
    const double c16 = c8 + 0.5 * c9;
    const double c17 = sin(c1);
    const double c18 = sin(c3);
    const double c19 = -c17 * c6 + 3.8312105531582844 * c18;
    const double c20 = c16 * c19;
    const double c21 = 0.25 * c13 + 0.5;
    const double c22 = -3.1415926535897932 * c0;
    const double c23 = c21 * sin(c22);
    const double c24 = cos(c12);
    const double c25 = 9.8696044010893586;
    const double c26 = c14 * c25;
    const double c27 = 0.00297441998810232 * c26 * (c8 - 2.0 * c9);
    const double c28 = 1.0 / c25;
    const double c29 = c14 * c14;
    const double c30 = 5.9488399762046401 * c25;
    const double c31 = c5 * (-2.97441998810232 * c25 * c9 - c30 * c8);
    const double c32 = 0.1681 * c28 * c29;
    const double c33 = c16 * (-c2 * c30 + 5.9488399762046401 * c25 * c4);
    const double c34 = c10 * c15 * c19;
    r.u = c11 * c15;
    r.v = -c15 * c20;
    r.p = c23 * c9;
    r.fx = 0.1681 * (c10 * c10) * c19 * c28 * c29 * c5 + 2.05 * c10 * c24 * c5 - c20 * c31 * c32 -
           c21 * c6 * c9 * cos(c22) - c27 * c5 - c27 * (c2 - c4);
    r.fy = c10 * c16 * (c19 * c19) * c32 - c11 * c32 * c33 + 0.0059488399762046401 * c16 * c26 * (c17 - 2.0 * c18) -
           2.05 * c20 * c24 - c23 * c6 * c8 + 0.00148720999405116 * c26 * (2.0 * c17 - c18) * (2.0 * c8 + c9);
    r.ux = c34;
    r.uy = c15 * c31;
    r.vx = -c15 * c33;
    r.vy = -c34;
    return r;
}

// ---- momentum system -----------------------------------------------------------------------
// One thread per row i. Element terms as in src/cfa/fem_kernels.cuh:89-180; Dirichlet rows as
// in :431-453.
// Row assembly kernels accumulate a row's matrix values in per-thread LDS slots (one wavefront
// per workgroup) instead of read-modify-writes in global memory; the sums are the same, from 0
// in ascending element order. Rows longer than ROW_SLOTS accumulate in place.
constexpr int ROW_SLOTS = 16;
constexpr int ROWS_PER_BLOCK = 64;

__global__ void __launch_bounds__(ROWS_PER_BLOCK)
k_momentum_rows(int n, const int* __restrict__ ev, const double* __restrict__ x, const double* __restrict__ y,
                const int* __restrict__ tag, const int* __restrict__ rp, const int* __restrict__ ci,
                const int* __restrict__ inc_ptr, const int* __restrict__ inc_elem, const int* __restrict__ inc_pos,
                const double* __restrict__ u_n, const double* __restrict__ v_n, const double* __restrict__ u_prev,
                const double* __restrict__ v_prev, const double* __restrict__ p, double nu, double rho, double a0,
                double a1, double a2, double c1, double c2, double tau_time, double skew,
                const double* __restrict__ fx, const double* __restrict__ fy, double u_max, double mms_time,
                OutletMode om, double* __restrict__ A, double* __restrict__ rhs_u, double* __restrict__ rhs_v,
                double* __restrict__ mass, double* __restrict__ u, double* __restrict__ v) {
    __shared__ double acc[ROW_SLOTS][ROWS_PER_BLOCK];
    const int lane = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + lane;
    if (i >= n) return;
    const int t = tag[i];
    const bool dirichlet = (t == BC_INLET || t == BC_WALL || t == BC_CYLINDER || (t == BC_OUTLET && om.dirichlet));
    const int rs = rp[i], re = rp[i + 1];
    const bool in_lds = re - rs <= ROW_SLOTS;
    if (!dirichlet)
        for (int k = rs; k < re; ++k) {
            if (in_lds) acc[k - rs][lane] = 0.0;
            else A[k] = 0.0;
        }
    double fu_acc = 0.0, fv_acc = 0.0, m_acc = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        const double area = g.area;
        m_acc += area / 3.0;
        if (dirichlet) continue;
        double uh[3], vh[3], wu_bar = 0.0, wv_bar = 0.0, div_w = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double un = u_n[nd[k]], vn = v_n[nd[k]], um = u_prev[nd[k]], vm = v_prev[nd[k]];
            uh[k] = a1 * un + a2 * um;
            vh[k] = a1 * vn + a2 * vm;
            const double wu = c1 * un + c2 * um, wv = c1 * vn + c2 * vm;
            wu_bar += wu / 3.0;
            wv_bar += wv / 3.0;
            div_w += wu * g.dNdx[k] + wv * g.dNdy[k];
        }
        const double vel = sqrt(wu_bar * wu_bar + wv_bar * wv_bar) + 1.0e-12;
        const double hK = sqrt(2.0 * area);
        const double tt = 2.0 / tau_time, tc = 2.0 * vel / hK, td = 4.0 * nu / (hK * hK);
        const double tau = 1.0 / sqrt(tt * tt + tc * tc + 9.0 * td * td);
        const double p0 = p[nd[0]], p1 = p[nd[1]], p2 = p[nd[2]];
        const double dpdx = (p0 * g.dNdx[0] + p1 * g.dNdx[1] + p2 * g.dNdx[2]) / rho;
        const double dpdy = (p0 * g.dNdy[0] + p1 * g.dNdy[1] + p2 * g.dNdy[2]) / rho;
        double ug[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) ug[k] = wu_bar * g.dNdx[k] + wv_bar * g.dNdy[k];
        // the owner's values, selected without a dynamic index
        const double dxa = a == 0 ? g.dNdx[0] : a == 1 ? g.dNdx[1] : g.dNdx[2];
        const double dya = a == 0 ? g.dNdy[0] : a == 1 ? g.dNdy[1] : g.dNdy[2];
        const double uga = a == 0 ? ug[0] : a == 1 ? ug[1] : ug[2];
        double f_u = 0.0, f_v = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double M = (a == b) ? (area / 6.0) : (area / 12.0);
            const double K = area * (dxa * g.dNdx[b] + dya * g.dNdy[b]);
            const double C = (area / 3.0) * ug[b] + skew * div_w * M;
            const double S = area * tau * uga * ug[b];
            const double Ms = (area / 3.0) * tau * uga;
            const double Aab = a0 * (M + Ms) + nu * K + C + S;
            const int pos = inc_pos[3 * s + b];
            if (pos >= 0) {
                if (in_lds) acc[pos - rs][lane] += Aab;
                else A[pos] += Aab;
            }
            f_u += (M + Ms) * uh[b];
            f_v += (M + Ms) * vh[b];
            if (fx) {
                f_u += (M + Ms) * fx[nd[b]];
                f_v += (M + Ms) * fy[nd[b]];
            }
        }
        const double tw = (area / 3.0) + area * tau * uga;
        f_u -= tw * dpdx;
        f_v -= tw * dpdy;
        fu_acc += f_u;
        fv_acc += f_v;
    }
    mass[i] = m_acc;
    if (dirichlet) {
        double ub = 0.0, vb = 0.0;
        if (mms_time >= 0.0) {
            const Mms r = mms_exact(x[i], y[i], mms_time);
            ub = r.u;
            vb = r.v;
        } else if (t == BC_INLET || t == BC_OUTLET) {   // the outlet only with om.dirichlet
            const double H = 0.41, yi = y[i];
            ub = 4.0 * u_max * yi * (H - yi) / (H * H) * (t == BC_OUTLET ? om.scale : 1.0);
        }
        for (int k = rs; k < re; ++k) A[k] = (ci[k] == i) ? 1.0 : 0.0;
        rhs_u[i] = ub;  rhs_v[i] = vb;
        u[i] = ub;      v[i] = vb;
    } else {
        if (in_lds)
            for (int k = rs; k < re; ++k) A[k] = acc[k - rs][lane];
        rhs_u[i] = fu_acc;
        rhs_v[i] = fv_acc;
    }
}

__global__ void k_mms_forcing(int n, const double* __restrict__ x, const double* __restrict__ y, double t,
                              double* __restrict__ fx, double* __restrict__ fy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Mms r = mms_exact(x[i], y[i], t);
    fx[i] = r.fx;
    fy[i] = r.fy;
}

// ---- pressure system -------------------------------------------------------------------------
__global__ void k_oss_weights(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                              const double* __restrict__ y, const double* __restrict__ us,
                              const double* __restrict__ vs, double dt, double rho, double nu, double stab,
                              double* __restrict__ s) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
    const double area = 0.5 * fabs((x[n1] - x[n0]) * (y[n2] - y[n0]) - (x[n2] - x[n0]) * (y[n1] - y[n0]));
    const double h = sqrt(2.0 * area);
    const double ub = (us[n0] + us[n1] + us[n2]) / 3.0;
    const double vb = (vs[n0] + vs[n1] + vs[n2]) / 3.0;
    const double tau = stab / (rho * (4.0 * nu / (h * h) + 2.0 * sqrt(ub * ub + vb * vb) / h) + 1.0e-300);
    s[e] = rho * tau / dt;
}

// G(f) row gather (src/cfa/fem_kernels.cuh:726-757)
__global__ void k_scalar_gradient_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                       const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                       const int* __restrict__ inc_elem, const double* __restrict__ f,
                                       const double* __restrict__ s_el, double* __restrict__ gx,
                                       double* __restrict__ gy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0, ay = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int e = inc_elem[s] / 3;
        const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
        const double x0 = x[n0], y0 = y[n0], x1 = x[n1], y1 = y[n1], x2 = x[n2], y2 = y[n2];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double area = 0.5 * fabs(detJ);
        if (area < 1.0e-14) continue;
        const double inv2A = 1.0 / (2.0 * area);
        const double g_x = (f[n0] * (y1 - y2) + f[n1] * (y2 - y0) + f[n2] * (y0 - y1)) * inv2A;
        const double g_y = (f[n0] * (x2 - x1) + f[n1] * (x0 - x2) + f[n2] * (x1 - x0)) * inv2A;
        const double w = (s_el ? sqrt(s_el[e]) : 1.0) * area / 3.0;
        ax += w * g_x;
        ay += w * g_y;
    }
    gx[i] = ax;
    gy[i] = ay;
}

// Lumped mass row gather (src/cfa/fem_kernels.cuh:890-909)
__global__ void k_lumped_mass_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                   const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                   const int* __restrict__ inc_elem, double* __restrict__ mass) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int e = inc_elem[s] / 3;
        const int n0 = ev[3 * e], n1 = ev[3 * e + 1], n2 = ev[3 * e + 2];
        const double detJ = (x[n1] - x[n0]) * (y[n2] - y[n0]) - (x[n2] - x[n0]) * (y[n1] - y[n0]);
        m += fabs(detJ) / 6.0;
    }
    mass[i] = m;
}

// PPE rows with OSS operator and outlet Dirichlet rows (src/cfa/fem_kernels.cuh:201-282, 460-490,
// 807-852).
__global__ void __launch_bounds__(ROWS_PER_BLOCK)
k_pressure_rows(int n, int nnz, const int* __restrict__ ev, const double* __restrict__ x,
                const double* __restrict__ y, const int* __restrict__ tag, const int* __restrict__ rp,
                const int* __restrict__ ci, const int* __restrict__ inc_ptr, const int* __restrict__ inc_elem,
                const int* __restrict__ inc_pos, const double* __restrict__ us, const double* __restrict__ vs,
                const double* __restrict__ p, const double* __restrict__ gpx, const double* __restrict__ gpy,
                const double* __restrict__ mass, const double* __restrict__ s_el, double dt, double rho,
                double stab, double* __restrict__ K, double* __restrict__ rhs, double* __restrict__ div_num,
                double* __restrict__ dp, double* __restrict__ oss_val, double* __restrict__ inv_m, OutletMode om) {
    // the row's values (K, Bx, By, BTx, BTy) accumulate in this thread's LDS slots, from 0 in
    // ascending element order as before; rows longer than ROW_SLOTS accumulate in place
    __shared__ double acc[5][ROW_SLOTS][ROWS_PER_BLOCK];
    const int t = threadIdx.x;
    const int i = blockIdx.x * blockDim.x + t;
    if (i >= n) return;
    const bool dirichlet = pressure_dirichlet(i, tag[i], om);
    const int rs = rp[i], re = rp[i + 1];
    const bool in_lds = re - rs <= ROW_SLOTS;
    double* Bx = oss_val;
    double* By = oss_val ? oss_val + nnz : nullptr;
    double* BTx = oss_val ? oss_val + 2 * (size_t)nnz : nullptr;
    double* BTy = oss_val ? oss_val + 3 * (size_t)nnz : nullptr;
    const int n_arr = oss_val ? 5 : 1;
    for (int k = rs; k < re; ++k) {
        if (in_lds) {
            for (int f = 0; f < n_arr; ++f) acc[f][k - rs][t] = 0.0;
        } else {
            K[k] = 0.0;
            if (oss_val) { Bx[k] = 0.0; By[k] = 0.0; BTx[k] = 0.0; BTy[k] = 0.0; }
        }
    }
    auto add = [&](int f, int pos, double v) {
        if (in_lds) {
            acc[f][pos - rs][t] += v;
            return;
        }
        double* dst = f == 0 ? K : f == 1 ? Bx : f == 2 ? By : f == 3 ? BTx : BTy;
        dst[pos] += v;
    };
    double r_acc = 0.0, d_acc = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        const double area = g.area;
        // the owner's shape-function gradient, selected without a dynamic index into g
        const double dxa = a == 0 ? g.dNdx[0] : a == 1 ? g.dNdx[1] : g.dNdx[2];
        const double dya = a == 0 ? g.dNdy[0] : a == 1 ? g.dNdy[1] : g.dNdy[2];
        const double div_u = us[nd[0]] * g.dNdx[0] + us[nd[1]] * g.dNdx[1] + us[nd[2]] * g.dNdx[2] +
                             vs[nd[0]] * g.dNdy[0] + vs[nd[1]] * g.dNdy[1] + vs[nd[2]] * g.dNdy[2];
        const double rc = -(rho / dt) * (area / 3.0) * div_u;
        double sk = 0.0, sx = 0.0, sy = 0.0, gxp = 0.0, gyp = 0.0;
        if (stab > 0.0) {
            sk = s_el[e];
            const double rsk = sqrt(sk);
            gxp = p[nd[0]] * g.dNdx[0] + p[nd[1]] * g.dNdx[1] + p[nd[2]] * g.dNdx[2];
            gyp = p[nd[0]] * g.dNdy[0] + p[nd[1]] * g.dNdy[1] + p[nd[2]] * g.dNdy[2];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double im = 1.0 / fmax(mass[nd[k]], 1.0e-14);
                sx += gpx[nd[k]] * im;
                sy += gpy[nd[k]] * im;
            }
            sx *= rsk / 3.0;
            sy *= rsk / 3.0;
        }
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double Kab = area * (dxa * g.dNdx[b] + dya * g.dNdy[b]);
            const int pos = inc_pos[3 * s + b];
            if (pos >= 0) add(0, pos, (1.0 + sk) * Kab);
        }
        const double stab_a = sk * area * (dxa * gxp + dya * gyp) - area * (dxa * sx + dya * sy);
        r_acc += rc - stab_a;
        d_acc += (area / 3.0) * div_u;
        if (oss_val) {
            const double w = sqrt(s_el[e]) * area / 3.0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int pos = inc_pos[3 * s + b];
                if (pos < 0) continue;
                add(1, pos, w * g.dNdx[b]);
                add(2, pos, w * g.dNdy[b]);
                add(3, pos, w * dxa);
                add(4, pos, w * dya);
            }
        }
    }
    if (in_lds)
        for (int k = rs; k < re; ++k) {
            K[k] = acc[0][k - rs][t];
            if (oss_val) {
                Bx[k] = acc[1][k - rs][t];
                By[k] = acc[2][k - rs][t];
                BTx[k] = acc[3][k - rs][t];
                BTy[k] = acc[4][k - rs][t];
            }
        }
    div_num[i] = d_acc;
    if (oss_val) inv_m[i] = 1.0 / fmax(mass[i], 1.0e-14);
    if (dirichlet) {
        for (int k = rs; k < re; ++k) K[k] = (ci[k] == i) ? 1.0 : 0.0;
        // a pinned row keeps its assembled value for the mean removal, which then zeroes it
        rhs[i] = om.dirichlet ? r_acc : 0.0;
        dp[i] = 0.0;
    } else {
        for (int k = rs; k < re; ++k)
            if (ci[k] != i && pressure_dirichlet(ci[k], tag[ci[k]], om)) K[k] = 0.0;
        rhs[i] = r_acc;
    }
}

// OSS operator rows only (src/cfa/fem_kernels.cuh:807-852)
__global__ void k_oss_operator_rows(int n, int nnz, const int* __restrict__ ev, const double* __restrict__ x,
                                    const double* __restrict__ y, const int* __restrict__ rp,
                                    const int* __restrict__ inc_ptr, const int* __restrict__ inc_elem,
                                    const int* __restrict__ inc_pos, const double* __restrict__ s_el,
                                    const double* __restrict__ mass, double* __restrict__ oss_val,
                                    double* __restrict__ inv_m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* Bx = oss_val;
    double* By = oss_val + nnz;
    double* BTx = oss_val + 2 * (size_t)nnz;
    double* BTy = oss_val + 3 * (size_t)nnz;
    for (int k = rp[i]; k < rp[i + 1]; ++k) { Bx[k] = 0.0; By[k] = 0.0; BTx[k] = 0.0; BTy[k] = 0.0; }
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        Geo g;
        if (!geometry(x, y, ev[3 * e], ev[3 * e + 1], ev[3 * e + 2], g)) continue;
        const double w = sqrt(s_el[e]) * g.area / 3.0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int pos = inc_pos[3 * s + b];
            if (pos < 0) continue;
            Bx[pos] += w * g.dNdx[b];
            By[pos] += w * g.dNdy[b];
            BTx[pos] += w * g.dNdx[a];
            BTy[pos] += w * g.dNdy[a];
        }
    }
    inv_m[i] = 1.0 / fmax(mass[i], 1.0e-14);
}

// out -= B^T M^-1 B in: stage 1 g = M^-1 B in (all rows), stage 2 on non-Dirichlet rows
// (src/cfa/fem_kernels.cuh:855-885)
template <int TPR>
__global__ void k_oss_B(int n, int nnz, const int* __restrict__ rp, const int* __restrict__ ci,
                        const double* __restrict__ oss, const double* __restrict__ inv_m,
                        const double* __restrict__ in, double* __restrict__ gx, double* __restrict__ gy) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = gtid / TPR, lane = gtid % TPR;
    double sx = 0.0, sy = 0.0;
    if (k < n)
        row_entries_n<TPR>(ci, RowValues<2, double>{{oss, oss + (size_t)nnz}}, rp[k], rp[k + 1], lane,
                           [&](const double* v, int j) {
                               const double xj = in[j];
                               sx += v[0] * xj;
                               sy += v[1] * xj;
                           });
    sx = group_sum<TPR>(sx);
    sy = group_sum<TPR>(sy);
    if (k < n && lane == 0) {
        gx[k] = sx * inv_m[k];
        gy[k] = sy * inv_m[k];
    }
}

template <int TPR>
__global__ void k_oss_BT(int n, int nnz, const int* __restrict__ rp, const int* __restrict__ ci,
                         const double* __restrict__ oss, const int* __restrict__ tag, OutletMode om,
                         const double* __restrict__ gx, const double* __restrict__ gy, double* __restrict__ out) {
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = gtid / TPR, lane = gtid % TPR;
    const double* BTx = oss + 2 * (size_t)nnz;
    const double* BTy = oss + 3 * (size_t)nnz;
    double s = 0.0;
    const bool active = i < n && !pressure_dirichlet(i, tag[i], om);
    if (active)
        row_entries_n<TPR>(ci, RowValues<2, double>{{BTx, BTy}}, rp[i], rp[i + 1], lane,
                           [&](const double* v, int k) { s += v[0] * gx[k] + v[1] * gy[k]; });
    s = group_sum<TPR>(s);
    if (active && lane == 0) out[i] -= s;
}

// Matrix-free OSS subtraction (src/cfa/fem_kernels.cuh:763-797), row gather
__global__ void k_oss_matrix_free_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                       const double* __restrict__ y, const int* __restrict__ tag, OutletMode om,
                                       const int* __restrict__ inc_ptr, const int* __restrict__ inc_elem,
                                       const double* __restrict__ s_el, const double* __restrict__ gx,
                                       const double* __restrict__ gy, const double* __restrict__ mass,
                                       double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || pressure_dirichlet(i, tag[i], om)) return;
    double acc = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        double sx = 0.0, sy = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double im = 1.0 / fmax(mass[nd[k]], 1.0e-14);
            sx += gx[nd[k]] * im;
            sy += gy[nd[k]] * im;
        }
        const double w = sqrt(s_el[e]) * g.area / 3.0;
        acc += -w * (g.dNdx[a] * sx + g.dNdy[a] * sy);
    }
    out[i] += acc;
}

// out = c K f (src/cfa/fem_kernels.cuh:627-649)
__global__ void k_stiffness_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                 const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                 const int* __restrict__ inc_elem, const double* __restrict__ f, double c,
                                 double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        const double gx = f[nd[0]] * g.dNdx[0] + f[nd[1]] * g.dNdx[1] + f[nd[2]] * g.dNdx[2];
        const double gy = f[nd[0]] * g.dNdy[0] + f[nd[1]] * g.dNdy[1] + f[nd[2]] * g.dNdy[2];
        acc += c * g.area * (g.dNdx[a] * gx + g.dNdy[a] * gy);
    }
    out[i] = acc;
}

// ---- correction ------------------------------------------------------------------------------
// Nodal sums of |K|/3 grad dp and |K|/3 omega_h (src/cfa/fem_kernels.cuh:536-582)
__global__ void k_grad_dp_vort_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                    const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                    const int* __restrict__ inc_elem, const double* __restrict__ dp,
                                    const double* __restrict__ u, const double* __restrict__ v,
                                    double* __restrict__ gx, double* __restrict__ gy, double* __restrict__ vn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0, ay = 0.0, av = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int e = inc_elem[s] / 3;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        const double dpdx = dp[nd[0]] * g.dNdx[0] + dp[nd[1]] * g.dNdx[1] + dp[nd[2]] * g.dNdx[2];
        const double dpdy = dp[nd[0]] * g.dNdy[0] + dp[nd[1]] * g.dNdy[1] + dp[nd[2]] * g.dNdy[2];
        const double dvdx = v[nd[0]] * g.dNdx[0] + v[nd[1]] * g.dNdx[1] + v[nd[2]] * g.dNdx[2];
        const double dudy = u[nd[0]] * g.dNdy[0] + u[nd[1]] * g.dNdy[1] + u[nd[2]] * g.dNdy[2];
        const double om = dvdx - dudy;
        const double w = g.area / 3.0;
        ax += w * dpdx;
        ay += w * dpdy;
        av += w * om;
    }
    gx[i] = ax;
    gy[i] = ay;
    vn[i] = av;
}

// src/cfa/fem_kernels.cuh:592-622 (default outlet: outlet velocities are corrected)
__global__ void k_correction(int n, const int* __restrict__ tag, int outlet_dirichlet, double dt, double rho,
                             double p_update,
                             double rot, const double* __restrict__ div_num, const double* __restrict__ mass,
                             const double* __restrict__ gx, const double* __restrict__ gy,
                             const double* __restrict__ vn, const double* __restrict__ dp, double* __restrict__ u,
                             double* __restrict__ v, double* __restrict__ p, double* __restrict__ vort) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = fmax(mass[i], 1.0e-14);
    vort[i] = vn[i] / m;
    p[i] += p_update * (dp[i] - rot * div_num[i] / m);
    const int t = tag[i];
    if (t == BC_INTERIOR || (t == BC_OUTLET && !outlet_dirichlet)) {
        u[i] -= (dt / rho) * (gx[i] / m);
        v[i] -= (dt / rho) * (gy[i] / m);
    }
}

// ---- pure-Neumann PPE (outlet Dirichlet) --------------------------------------------------------
constexpr int MEAN_BLOCK = 256;

// Block sums of f and m over MEAN_BLOCK entries with the in-block tree of
// src/cfa/fem_kernels.cuh:493-503 k_sum (s[t] += s[t + w], w = 128, 64, ..., 1).
// part[b] = sum of f over block b, part[nb + b] = sum of m.
__global__ void __launch_bounds__(MEAN_BLOCK)
k_mean_block_sums(int n, const double* __restrict__ f, const double* __restrict__ m, double* __restrict__ part) {
    __shared__ double s[2][MEAN_BLOCK];
    const int t = threadIdx.x;
    const int i = blockIdx.x * MEAN_BLOCK + t;
    s[0][t] = i < n ? f[i] : 0.0;
    s[1][t] = i < n ? m[i] : 0.0;
    __syncthreads();
    for (int w = MEAN_BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            s[0][t] += s[0][t + w];
            s[1][t] += s[1][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[blockIdx.x] = s[0][0];
        part[gridDim.x + blockIdx.x] = s[1][0];
    }
}

// sums[q] = 0 + part[q nb] + part[q nb + 1] + ... in ascending block order (thread q)
__global__ void k_mean_final(int nb, const double* __restrict__ part, double* __restrict__ sums) {
    const int q = threadIdx.x;
    if (q >= 2) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += part[(size_t)q * nb + b];
    sums[q] = acc;
}

// f_i -= (sum f / sum m) m_i (src/cfa/fem_kernels.cuh:506-510 k_subtract_mean), then the pinned
// row's rhs = 0 (:475-479 k_apply_bc_pressure)
__global__ void k_subtract_mean(int n, const double* __restrict__ sums, const double* __restrict__ m, int pin,
                                double* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f[i] = i == pin ? 0.0 : f[i] - sums[0] / sums[1] * m[i];
}

// ---- forces ------------------------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024;

// Volume-integral force bracket of one element (src/cfa/fem_kernels.cuh:1046-1123)
__device__ void element_force(int e, const int* __restrict__ ev, const int* __restrict__ tag,
                              const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ u, const double* __restrict__ v,
                              const double* __restrict__ uo, const double* __restrict__ vo,
                              const double* __restrict__ up, const double* __restrict__ vp,
                              const double* __restrict__ p, double b0, double b1, double b2, double nu,
                              double rho, double& fx_out, double& fy_out) {
    fx_out = 0.0;
    fy_out = 0.0;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double phi[3];
    bool touches = false;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        phi[i] = (tag[nd[i]] == BC_CYLINDER) ? 1.0 : 0.0;
        touches = touches || phi[i] > 0.0;
    }
    if (!touches) return;
    Geo g;
    if (!geometry(x, y, nd[0], nd[1], nd[2], g)) return;
    double ue[3], ve[3], ute[3], vte[3];
    double dudx = 0, dudy = 0, dvdx = 0, dvdy = 0, dphidx = 0, dphidy = 0, pbar = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ue[i] = u[nd[i]];
        ve[i] = v[nd[i]];
        ute[i] = b0 * ue[i] + b1 * uo[nd[i]] + b2 * up[nd[i]];
        vte[i] = b0 * ve[i] + b1 * vo[nd[i]] + b2 * vp[nd[i]];
        dudx += ue[i] * g.dNdx[i];  dudy += ue[i] * g.dNdy[i];
        dvdx += ve[i] * g.dNdx[i];  dvdy += ve[i] * g.dNdy[i];
        dphidx += phi[i] * g.dNdx[i];
        dphidy += phi[i] * g.dNdy[i];
        pbar += p[nd[i]] / 3.0;
    }
    double fx = 0.0, fy = 0.0;
#pragma unroll
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
    fx_out = fx;
    fy_out = fy;
}

// Block partials: ws[4 * b + 0..1] volume, ws[4 * b + 2..3] surface
__global__ void __launch_bounds__(BLOCK)
k_forces_partial(int count, const int* __restrict__ layer, const int* __restrict__ ev,
                 const int* __restrict__ tag, const double* __restrict__ x, const double* __restrict__ y,
                 const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ uo,
                 const double* __restrict__ vo, const double* __restrict__ up, const double* __restrict__ vp,
                 const double* __restrict__ p, double b0, double b1, double b2, double nu, double rho,
                 int n_cyl, const int* __restrict__ edges, double* __restrict__ ws) {
    __shared__ double lds[BLOCK / WAVE];
    double fx = 0.0, fy = 0.0, sx = 0.0, sy = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const int e = layer ? layer[k] : k;
        double ex, ey;
        element_force(e, ev, tag, x, y, u, v, uo, vo, up, vp, p, b0, b1, b2, nu, rho, ex, ey);
        fx += ex;
        fy += ey;
    }
    // surface traction (src/cfa/fem_kernels.cuh:657-692)
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_cyl; k += gridDim.x * blockDim.x) {
        const int a = edges[3 * k], b = edges[3 * k + 1], e = edges[3 * k + 2];
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        const double x0 = x[nd[0]], y0 = y[nd[0]], x1 = x[nd[1]], y1 = y[nd[1]], x2 = x[nd[2]], y2 = y[nd[2]];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double dNdx[3] = {(y1 - y2) / detJ, (y2 - y0) / detJ, (y0 - y1) / detJ};
        const double dNdy[3] = {(x2 - x1) / detJ, (x0 - x2) / detJ, (x1 - x0) / detJ};
        double ux = 0, uy = 0, vx = 0, vy = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ux += u[nd[i]] * dNdx[i];  uy += u[nd[i]] * dNdy[i];
            vx += v[nd[i]] * dNdx[i];  vy += v[nd[i]] * dNdy[i];
        }
        const double dx = x[b] - x[a], dy = y[b] - y[a];
        const double len = sqrt(dx * dx + dy * dy);
        double nx = dy / len, ny = -dx / len;
        if (nx * (0.5 * (x[a] + x[b]) - 0.2) + ny * (0.5 * (y[a] + y[b]) - 0.2) < 0.0) { nx = -nx; ny = -ny; }
        const double pm = 0.5 * (p[a] + p[b]) / rho;
        sx += len * (nu * (ux * nx + uy * ny) - pm * nx);
        sy += len * (nu * (vx * nx + vy * ny) - pm * ny);
    }
    fx = block_sum(fx, lds);
    fy = block_sum(fy, lds);
    sx = block_sum(sx, lds);
    sy = block_sum(sy, lds);
    if (threadIdx.x == 0) {
        ws[4 * blockIdx.x + 0] = fx;
        ws[4 * blockIdx.x + 1] = fy;
        ws[4 * blockIdx.x + 2] = sx;
        ws[4 * blockIdx.x + 3] = sy;
    }
}

__global__ void k_forces_final(int nb, const double* __restrict__ ws, const double* __restrict__ p, int a0,
                               int b0, double w0, int a1, int b1, double w1, double* __restrict__ diag) {
    __shared__ double lds[BLOCK / WAVE];
    double r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = sum_partials(ws + q, nb, 4, lds);
    if (threadIdx.x == 0) {
        diag[0] = r[0];
        diag[1] = r[1];
        diag[2] = (1.0 - w0) * p[a0] + w0 * p[b0];
        diag[3] = (1.0 - w1) * p[a1] + w1 * p[b1];
        diag[4] = r[2];
        diag[5] = r[3];
    }
}

// ---- indicators ----------------------------------------------------------------------------------
// Lumped mass and ZZ-recovered disturbance gradient (src/cfa/fem_kernels.cuh:921-972)
__global__ void k_recovered_gradient_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                          const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                          const int* __restrict__ inc_elem, const double* __restrict__ u,
                                          const double* __restrict__ v, double u_max, double* __restrict__ mass,
                                          double* __restrict__ grad_rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double H = 0.41;
    double m = 0.0, acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int e = inc_elem[s] / 3;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        const double detJ = (x[nd[1]] - x[nd[0]]) * (y[nd[2]] - y[nd[0]]) - (x[nd[2]] - x[nd[0]]) * (y[nd[1]] - y[nd[0]]);
        m += fabs(detJ) / 6.0;
        Geo g;
        if (!geometry(x, y, nd[0], nd[1], nd[2], g)) continue;
        double ud[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double yk = y[nd[k]];
            ud[k] = u[nd[k]] - 4.0 * u_max * yk * (H - yk) / (H * H);
        }
        const double gr[4] = {ud[0] * g.dNdx[0] + ud[1] * g.dNdx[1] + ud[2] * g.dNdx[2],
                              ud[0] * g.dNdy[0] + ud[1] * g.dNdy[1] + ud[2] * g.dNdy[2],
                              v[nd[0]] * g.dNdx[0] + v[nd[1]] * g.dNdx[1] + v[nd[2]] * g.dNdx[2],
                              v[nd[0]] * g.dNdy[0] + v[nd[1]] * g.dNdy[1] + v[nd[2]] * g.dNdy[2]};
        const double w = g.area / 3.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] += w * gr[c];
    }
    mass[i] = m;
#pragma unroll
    for (int c = 0; c < 4; ++c) grad_rec[(size_t)c * n + i] = acc[c];
}

// src/cfa/fem_kernels.cuh:995-1035
__global__ void k_indicator_elems(int ne, int n, const int* __restrict__ ev, const double* __restrict__ x,
                                  const double* __restrict__ y, const double* __restrict__ grad_rec,
                                  const double* __restrict__ mass, double* __restrict__ err) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double G[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double im = 1.0 / fmax(mass[nd[i]], 1.0e-14);
#pragma unroll
        for (int c = 0; c < 4; ++c) G[i][c] = grad_rec[(size_t)c * n + nd[i]] * im;
    }
    const double x0 = x[nd[0]], y0 = y[nd[0]];
    const double area = 0.5 * fabs((x[nd[1]] - x0) * (y[nd[2]] - y0) - (x[nd[2]] - x0) * (y[nd[1]] - y0));
    double omega = 0.0, eta_i = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        omega = fmax(omega, fabs(G[i][2] - G[i][1]));
        const int j = (i + 1) % 3;
        const double ex = x[nd[j]] - x[nd[i]], ey = y[nd[j]] - y[nd[i]];
        const double eu = (G[j][0] - G[i][0]) * ex + (G[j][1] - G[i][1]) * ey;
        const double evv = (G[j][2] - G[i][2]) * ex + (G[j][3] - G[i][3]) * ey;
        eta_i = fmax(eta_i, 0.125 * sqrt(eu * eu + evv * evv));
    }
    err[e] = sqrt(2.0 * area) * omega;
    err[ne + e] = eta_i;
}

// ---- manufactured-solution errors (src/cfa/fem_kernels.cuh:362-405) ------------------------------
__global__ void __launch_bounds__(BLOCK)
k_mms_errors_partial(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                     const double* __restrict__ y, const double* __restrict__ uh, const double* __restrict__ vh,
                     const double* __restrict__ ph, double t, double* __restrict__ ws) {
    __shared__ double lds[BLOCK / WAVE];
    const double a1 = 0.059715871789770, b1 = 0.470142064105115, w1 = 0.132394152788506;
    const double a2 = 0.797426985353087, b2 = 0.101286507323456, w2 = 0.125939180544827;
    const double Lq[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3}, {a1, b1, b1}, {b1, a1, b1}, {b1, b1, a1},
                             {a2, b2, b2}, {b2, a2, b2}, {b2, b2, a2}};
    const double W[7] = {0.225, w1, w1, w1, w2, w2, w2};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
        const int n[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        const double x0 = x[n[0]], y0 = y[n[0]], x1 = x[n[1]], y1 = y[n[1]], x2 = x[n[2]], y2 = y[n[2]];
        const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
        const double area = 0.5 * fabs(detJ);
        const double inv = 1.0 / detJ;
        const double dNdx[3] = {(y1 - y2) * inv, (y2 - y0) * inv, (y0 - y1) * inv};
        const double dNdy[3] = {(x2 - x1) * inv, (x0 - x2) * inv, (x1 - x0) * inv};
        double uhx = 0, uhy = 0, vhx = 0, vhy = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            uhx += uh[n[k]] * dNdx[k];  uhy += uh[n[k]] * dNdy[k];
            vhx += vh[n[k]] * dNdx[k];  vhy += vh[n[k]] * dNdy[k];
        }
        for (int q = 0; q < 7; ++q) {
            const double xq = Lq[q][0] * x0 + Lq[q][1] * x1 + Lq[q][2] * x2;
            const double yq = Lq[q][0] * y0 + Lq[q][1] * y1 + Lq[q][2] * y2;
            double uq = 0, vq = 0, pq = 0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                uq += Lq[q][k] * uh[n[k]];  vq += Lq[q][k] * vh[n[k]];  pq += Lq[q][k] * ph[n[k]];
            }
            const Mms r = mms_exact(xq, yq, t);
            const double w = W[q] * area;
            acc[0] += w * ((uq - r.u) * (uq - r.u) + (vq - r.v) * (vq - r.v));
            acc[1] += w * ((uhx - r.ux) * (uhx - r.ux) + (uhy - r.uy) * (uhy - r.uy) +
                           (vhx - r.vx) * (vhx - r.vx) + (vhy - r.vy) * (vhy - r.vy));
            acc[2] += w * (pq - r.p) * (pq - r.p);
            acc[3] += w * (r.u * r.u + r.v * r.v);
            acc[4] += w * (r.ux * r.ux + r.uy * r.uy + r.vx * r.vx + r.vy * r.vy);
            acc[5] += w * r.p * r.p;
        }
    }
    for (int k = 0; k < 6; ++k) {
        const double s = block_sum(acc[k], lds);
        if (threadIdx.x == 0) ws[6 * blockIdx.x + k] = s;
    }
}

__global__ void k_sum_partials6(int nb, const double* __restrict__ ws, double* __restrict__ out) {
    __shared__ double lds[BLOCK / WAVE];
    for (int k = 0; k < 6; ++k) {
        const double s = sum_partials(ws + k, nb, 6, lds);
        if (threadIdx.x == 0) out[k] = s;
    }
}

inline bool mesh_ok(const cfa_mesh_t* m) {
    return m && m->n_nodes >= 0 && m->n_elems >= 0 && m->elem_v && m->x && m->y && m->bc_tag &&
           m->row_ptr && m->col_ind && m->inc_ptr && m->inc_elem && m->inc_pos &&
           (!m->outlet_dirichlet || m->n_nodes == 0 || (m->p_pin >= 0 && m->p_pin < m->n_nodes));
}

inline OutletMode outlet_of(const cfa_mesh_t* m) {
    return OutletMode{m->outlet_dirichlet, m->p_pin, m->outlet_scale};
}

// rhs -= (sum rhs / sum mass) mass, rhs[pin] = 0, with the block sums and totals in ws
int remove_mean(const cfa_mesh_t* m, const double* mass, double* rhs, double* ws, cfa_stream_t stream) {
    const int n = m->n_nodes;
    const int nb = blocks_for(n, MEAN_BLOCK);
    hipLaunchKernelGGL(k_mean_block_sums, dim3(nb), dim3(MEAN_BLOCK), 0, as_stream(stream), n, rhs, mass, ws);
    hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(WAVE), 0, as_stream(stream), nb, ws, ws + 2 * (size_t)nb);
    hipLaunchKernelGGL(k_subtract_mean, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n,
                       ws + 2 * (size_t)nb, mass, m->p_pin, rhs);
    return launch_status();
}

}  // namespace

// ---- the mean removal without allocation -------------------------------------------------------
// Workspace of cfa_pressure_remove_mean_noalloc: the block sums of at most CFA_MEAN_NOALLOC_MAX_NODES
// nodes (two per block of MEAN_BLOCK) and the two totals, 512 KiB + 16 B, one copy per GPU held
// until the process ends. Host code never passes the array's name (on the host it is only a
// stand-in): it uses the device address hipGetSymbolAddress gives for the current device, taken
// once per device and kept in scratch_ptr.
constexpr size_t MEAN_NOALLOC_DOUBLES = 2 * (size_t)(CFA_MEAN_NOALLOC_MAX_NODES / 256) + 2;
static_assert(CFA_MEAN_NOALLOC_MAX_NODES % 256 == 0, "whole blocks of the mean removal");
static __device__ __align__(256) double scratch[MEAN_NOALLOC_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on the current device (looked up on the first request)
int scratch_address(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= MAX_DEVICES) return CFA_E_ARG;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = scratch_ptr[dev];
    return CFA_OK;
}
}  // namespace

namespace cfa_dev {
int scratch_setup() {
    double* p = nullptr;
    return scratch_address(&p);
}
}  // namespace cfa_dev

extern "C" {

int cfa_assemble_momentum(const cfa_mesh_t* m, const double* u_n, const double* v_n,
                          const double* u_prev, const double* v_prev, const double* p, double nu,
                          double rho, double a0, double a1, double a2, double c1, double c2,
                          double tau_time, double skew, const double* fx, const double* fy,
                          double u_max, double mms_time, double* A_val, double* rhs_u,
                          double* rhs_v, double* mass, double* u, double* v, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u_n || !v_n || !u_prev || !v_prev || !p || !A_val || !rhs_u || !rhs_v || !mass || !u || !v)
        return CFA_E_ARG;
    if (u_n == u || v_n == v || ((fx == nullptr) != (fy == nullptr))) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_momentum_rows, dim3(blocks_for(m->n_nodes, ROWS_PER_BLOCK)), dim3(ROWS_PER_BLOCK), 0,
                       as_stream(stream),
                       m->n_nodes, m->elem_v, m->x, m->y, m->bc_tag, m->row_ptr, m->col_ind, m->inc_ptr,
                       m->inc_elem, m->inc_pos, u_n, v_n, u_prev, v_prev, p, nu, rho, a0, a1, a2, c1, c2,
                       tau_time, skew, fx, fy, u_max, mms_time, outlet_of(m), A_val, rhs_u, rhs_v, mass, u, v);
    return launch_status();
}

int cfa_mms_forcing(int n, const double* x, const double* y, double t, double* fx, double* fy,
                    cfa_stream_t stream) {
    if (n < 0 || !x || !y || !fx || !fy) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_mms_forcing, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n, x, y, t, fx, fy);
    return launch_status();
}

int cfa_oss_weights(const cfa_mesh_t* m, const double* u, const double* v, double dt, double rho,
                    double nu, double stab, double* elem_s, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u || !v || !elem_s) return CFA_E_ARG;
    if (m->n_elems == 0) return CFA_OK;
    hipLaunchKernelGGL(k_oss_weights, dim3(blocks_for(m->n_elems)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_elems, m->elem_v, m->x, m->y, u, v, dt, rho, nu, stab, elem_s);
    return launch_status();
}

int cfa_scalar_gradient(const cfa_mesh_t* m, const double* f, const double* elem_s, double* gx,
                        double* gy, cfa_stream_t stream) {
    if (!mesh_ok(m) || !f || !gx || !gy) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_scalar_gradient_rows, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_nodes, m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, f, elem_s, gx, gy);
    return launch_status();
}

int cfa_lumped_mass(const cfa_mesh_t* m, double* mass, cfa_stream_t stream) {
    if (!mesh_ok(m) || !mass) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_lumped_mass_rows, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_nodes, m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, mass);
    return launch_status();
}

int cfa_assemble_pressure(const cfa_mesh_t* m, const double* u_star, const double* v_star,
                          const double* p, const double* gpx, const double* gpy,
                          const double* mass, const double* elem_s, double dt, double rho,
                          double stab, double* Kp, double* rhs, double* div_num, double* dp,
                          double* oss_val, double* inv_m, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u_star || !v_star || !p || !mass || !Kp || !rhs || !div_num || !dp) return CFA_E_ARG;
    if (stab > 0.0 && (!gpx || !gpy || !elem_s)) return CFA_E_ARG;
    if (oss_val && (!inv_m || !elem_s)) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_pressure_rows, dim3(blocks_for(m->n_nodes, ROWS_PER_BLOCK)), dim3(ROWS_PER_BLOCK), 0,
                       as_stream(stream),
                       m->n_nodes, m->nnz, m->elem_v, m->x, m->y, m->bc_tag, m->row_ptr, m->col_ind,
                       m->inc_ptr, m->inc_elem, m->inc_pos, u_star, v_star, p, gpx, gpy, mass, elem_s, dt,
                       rho, stab, Kp, rhs, div_num, dp, oss_val, inv_m, outlet_of(m));
    return launch_status();
}

size_t cfa_pressure_mean_ws_doubles(int n) {
    return 2 * (size_t)blocks_for(n > 0 ? n : 1, MEAN_BLOCK) + 2;
}

int cfa_pressure_remove_mean(const cfa_mesh_t* m, const double* mass, double* rhs, double* ws,
                             cfa_stream_t stream) {
    if (!mesh_ok(m) || !m->outlet_dirichlet || !mass || !rhs || !ws) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    return remove_mean(m, mass, rhs, ws, stream);
}

int cfa_pressure_remove_mean_noalloc(const cfa_mesh_t* m, const double* mass, double* rhs, cfa_stream_t stream) {
    if (!mesh_ok(m) || !m->outlet_dirichlet || !mass || !rhs) return CFA_E_ARG;
    if (m->n_nodes > CFA_MEAN_NOALLOC_MAX_NODES) return CFA_E_SIZE;   // before any device work
    if (m->n_nodes == 0) return CFA_OK;
    double* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != CFA_OK) return rc;
    return remove_mean(m, mass, rhs, ws, stream);
}

int cfa_assemble_oss_operator(const cfa_mesh_t* m, const double* elem_s, const double* mass,
                              double* oss_val, double* inv_m, cfa_stream_t stream) {
    if (!mesh_ok(m) || !elem_s || !mass || !oss_val || !inv_m) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_oss_operator_rows, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_nodes, m->nnz, m->elem_v, m->x, m->y, m->row_ptr, m->inc_ptr, m->inc_elem, m->inc_pos,
                       elem_s, mass, oss_val, inv_m);
    return launch_status();
}

int cfa_oss_apply(const cfa_mesh_t* m, const double* oss_val, const double* inv_m,
                  const double* in, double* out, double* gx, double* gy, cfa_stream_t stream) {
    if (!mesh_ok(m) || !oss_val || !inv_m || !in || !out || !gx || !gy) return CFA_E_ARG;
    const int n = m->n_nodes;
    if (n == 0) return CFA_OK;
    constexpr int TPR = 4;
    const int blocks = blocks_for((long long)n * TPR);
    hipLaunchKernelGGL(k_oss_B<TPR>, dim3(blocks), dim3(BLOCK), 0, as_stream(stream), n, m->nnz, m->row_ptr,
                       m->col_ind, oss_val, inv_m, in, gx, gy);
    hipLaunchKernelGGL(k_oss_BT<TPR>, dim3(blocks), dim3(BLOCK), 0, as_stream(stream), n, m->nnz, m->row_ptr,
                       m->col_ind, oss_val, m->bc_tag, outlet_of(m), gx, gy, out);
    return launch_status();
}

int cfa_oss_apply_matrix_free(const cfa_mesh_t* m, const double* elem_s, const double* mass,
                              const double* in, double* out, double* gx, double* gy,
                              cfa_stream_t stream) {
    if (!mesh_ok(m) || !elem_s || !mass || !in || !out || !gx || !gy) return CFA_E_ARG;
    const int n = m->n_nodes;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_scalar_gradient_rows, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n,
                       m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, in, elem_s, gx, gy);
    hipLaunchKernelGGL(k_oss_matrix_free_rows, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n,
                       m->elem_v, m->x, m->y, m->bc_tag, outlet_of(m), m->inc_ptr, m->inc_elem, elem_s, gx, gy, mass,
                       out);
    return launch_status();
}

int cfa_stiffness_apply(const cfa_mesh_t* m, const double* f, double c, double* out,
                        cfa_stream_t stream) {
    if (!mesh_ok(m) || !f || !out || f == out) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_stiffness_rows, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_nodes, m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, f, c, out);
    return launch_status();
}

int cfa_correction(const cfa_mesh_t* m, double dt, double rho, double p_update, double rot,
                   const double* div_num, const double* mass, const double* dp, double* u,
                   double* v, double* p, double* vort, double* gx, double* gy, double* vort_num,
                   cfa_stream_t stream) {
    if (!mesh_ok(m) || !div_num || !mass || !dp || !u || !v || !p || !vort || !gx || !gy || !vort_num)
        return CFA_E_ARG;
    const int n = m->n_nodes;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_grad_dp_vort_rows, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n,
                       m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, dp, u, v, gx, gy, vort_num);
    hipLaunchKernelGGL(k_correction, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n, m->bc_tag,
                       m->outlet_dirichlet, dt, rho, p_update, rot, div_num, mass, gx, gy, vort_num, dp, u, v, p,
                       vort);
    return launch_status();
}

int cfa_forces(const cfa_mesh_t* m, const int* cyl_edges, int n_cyl, const int* layer,
               int n_layer, const double* u, const double* v, const double* u_old,
               const double* v_old, const double* u_prev, const double* v_prev, const double* p,
               double b0, double b1, double b2, double nu, double rho, const int* probe, double w0,
               double w1, double* diag, double* ws, cfa_stream_t stream) {
    if (!mesh_ok(m) || (n_cyl > 0 && !cyl_edges) || !u || !v || !u_old || !v_old || !u_prev || !v_prev || !p ||
        !probe || !diag || !ws || n_layer < 0)
        return CFA_E_ARG;
    const int count = layer ? n_layer : m->n_elems;
    const int work = count > n_cyl ? count : n_cyl;
    int nb = blocks_for(work > 0 ? work : 1);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    hipLaunchKernelGGL(k_forces_partial, dim3(nb), dim3(BLOCK), 0, as_stream(stream), count, layer, m->elem_v,
                       m->bc_tag, m->x, m->y, u, v, u_old, v_old, u_prev, v_prev, p, b0, b1, b2, nu, rho, n_cyl,
                       cyl_edges, ws);
    hipLaunchKernelGGL(k_forces_final, dim3(1), dim3(BLOCK), 0, as_stream(stream), nb, ws, p, probe[0], probe[1],
                       w0, probe[2], probe[3], w1, diag);
    return launch_status();
}

int cfa_error_indicator(const cfa_mesh_t* m, const double* u, const double* v, double u_max,
                        double* mass, double* grad_rec, double* elem_error, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u || !v || !mass || !grad_rec || !elem_error) return CFA_E_ARG;
    if (m->n_nodes == 0) return CFA_OK;
    hipLaunchKernelGGL(k_recovered_gradient_rows, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, as_stream(stream),
                       m->n_nodes, m->elem_v, m->x, m->y, m->inc_ptr, m->inc_elem, u, v, u_max, mass, grad_rec);
    if (m->n_elems > 0)
        hipLaunchKernelGGL(k_indicator_elems, dim3(blocks_for(m->n_elems)), dim3(BLOCK), 0, as_stream(stream),
                           m->n_elems, m->n_nodes, m->elem_v, m->x, m->y, grad_rec, mass, elem_error);
    return launch_status();
}

int cfa_mms_errors(const cfa_mesh_t* m, const double* u, const double* v, const double* p,
                   double t, double* out, double* ws, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u || !v || !p || !out || !ws) return CFA_E_ARG;
    int nb = blocks_for(m->n_elems > 0 ? m->n_elems : 1);
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    hipLaunchKernelGGL(k_mms_errors_partial, dim3(nb), dim3(BLOCK), 0, as_stream(stream), m->n_elems, m->elem_v,
                       m->x, m->y, u, v, p, t, ws);
    hipLaunchKernelGGL(k_sum_partials6, dim3(1), dim3(BLOCK), 0, as_stream(stream), nb, ws, out);
    return launch_status();
}

}  // extern "C"
