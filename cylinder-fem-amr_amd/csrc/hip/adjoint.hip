// Discrete adjoint of the BDF2 projection step for gfx950: vector-Jacobian products of the step's
// element assemblies and nodal updates. The forward step is the one of src/cylinder_fem_amr.cu:
// 277-538 (momentum fem_kernels.cuh:40-181 + :411-453, OSS weights :697-719, pressure :201-282
// + :460-531, OSS operator :763-885, rotational term :627-649, correction :536-622, volume forces
// :1046-1123); the reference has no adjoint.
//
// Each linear solve x = A(q)^{-1} b(q) of the step is differentiated through the scalar
//     Phi(q) = lambda^T (b(q) - A(q) x),   A^T lambda = x_bar,
// so q_bar = dPhi/dq with lambda and x held fixed. Phi is a sum of element terms. Each element's
// term is evaluated from its raw inputs (node coordinates, nodal velocities, viscosity) in
// forward-mode dual numbers, four directions per pass, so its gradient, including the
// coordinates' (shape sensitivities), comes from the element formulas themselves; inputs that
// enter linearly are differentiated by hand in the first pass. Couplings through the lumped
// mass and the OSS operator are linearised around fixed nodal fields so every term stays local
// to its element. Element contributions to nodes go to a per-(element, corner) array and are
// summed per node over the node -> element table in element order (no atomics; the result does
// not depend on scheduling).
#include "cfa_device.hpp"

using namespace cfa_dev;

namespace {

// Forward-mode dual number with N derivative components.
template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> dconst(double a) {
    Dual<N> r;
    r.v = a;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& b, double a) {
    return a * b;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
    Dual<N> r = b;
    r.v = a + b.v;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& b, double a) {
    return a + b;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& b, double a) {
    Dual<N> r = b;
    r.v = b.v - a;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v / b.v;
    const double inv = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
    return dconst<N>(a) / b;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
    return (1.0 / b) * a;
}
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dabs(const Dual<N>& a) {
    return a.v < 0.0 ? -a : a;
}

constexpr int ND = 4;   // derivative components per pass
using D = Dual<ND>;

struct Geo {
    double area, dNdx[3], dNdy[3];
};

// P1 gradients (src/cfa/fem_kernels.cuh:81-87); false for a degenerate element (|K| < 1e-14)
__device__ __forceinline__ bool geometry(const double* __restrict__ x, const double* __restrict__ y,
                                         const int* nd, Geo& g) {
    const double x0 = x[nd[0]], y0 = y[nd[0]], x1 = x[nd[1]], y1 = y[nd[1]], x2 = x[nd[2]], y2 = y[nd[2]];
    const double detJ = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    g.area = 0.5 * fabs(detJ);
    if (g.area < 1.0e-14) return false;
    const double inv2A = 1.0 / (2.0 * g.area);
    g.dNdx[0] = (y1 - y2) * inv2A;  g.dNdx[1] = (y2 - y0) * inv2A;  g.dNdx[2] = (y0 - y1) * inv2A;
    g.dNdy[0] = (x2 - x1) * inv2A;  g.dNdy[1] = (x0 - x2) * inv2A;  g.dNdy[2] = (x1 - x0) * inv2A;
    return true;
}

// the same gradients from dual coordinates X[3], Y[3]
struct GeoD {
    D area, dNdx[3], dNdy[3];
};
__device__ __forceinline__ bool geometry_d(const D* X, const D* Y, GeoD& g) {
    const D detJ = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
    g.area = 0.5 * dabs(detJ);
    if (g.area.v < 1.0e-14) return false;
    const D inv2A = 1.0 / (2.0 * g.area);
    g.dNdx[0] = (Y[1] - Y[2]) * inv2A;  g.dNdx[1] = (Y[2] - Y[0]) * inv2A;  g.dNdx[2] = (Y[0] - Y[1]) * inv2A;
    g.dNdy[0] = (X[2] - X[1]) * inv2A;  g.dNdy[1] = (X[0] - X[2]) * inv2A;  g.dNdy[2] = (X[1] - X[0]) * inv2A;
    return true;
}

__device__ __forceinline__ bool dirichlet_velocity(int t) {
    return t == BC_INLET || t == BC_WALL || t == BC_CYLINDER;
}

// Gradient of an element functional phi(in[0..n_in)) by forward passes of ND directions over the
// inputs from `first` on: grad[q] = d phi / d in[q]. `f(in, value_pass)` returns phi as a dual and
// reports a degenerate element by returning false in its ok flag.
template <typename F>
__device__ __forceinline__ bool element_gradient(int n_in, int first, const double* vals, double* grad, F f) {
    D in[32];
    for (int base = first; base < n_in; base += ND) {
        for (int q = 0; q < n_in; ++q) {
            in[q] = dconst<ND>(vals[q]);
            const int slot = q - base;
            if (slot >= 0 && slot < ND) in[q].d[slot] = 1.0;
        }
        bool ok = true;
        const D phi = f(in, base == first, ok);
        if (!ok) return false;
        for (int slot = 0; slot < ND; ++slot) {
            const int q = base + slot;
            if (q < n_in) grad[q] = phi.d[slot];
        }
    }
    return true;
}

// ---- momentum: Phi_K = sum over free rows a of lambda_a . (b_a^K - sum_b A_ab^K x_b) ------------
// inputs: X[3], Y[3] (coordinates), WU[3], WV[3] (advecting velocity c1 u^n + c2 u^{n-1}), nu;
// the right-hand-side data u^h, p enter linearly (their derivatives from the value pass).
// contrib fields (stride 3 ne): 0 u_n, 1 v_n, 2 u_{n-1}, 3 v_{n-1}, 4 p, 5 x, 6 y; nu_part[e]
__global__ void k_adj_momentum_elems(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                                     const double* __restrict__ y, const int* __restrict__ tag,
                                     const double* __restrict__ un, const double* __restrict__ vn,
                                     const double* __restrict__ um, const double* __restrict__ vm,
                                     const double* __restrict__ p, const double* __restrict__ us,
                                     const double* __restrict__ vs, const double* __restrict__ lu,
                                     const double* __restrict__ lv, double nu, double rho, double a0, double a1,
                                     double a2, double c1, double c2, double tau_time, double skew, int geom,
                                     double* __restrict__ contrib, double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double lam_u[3], lam_v[3], uh[3], vh[3], pe[3], xs_u[3], xs_v[3], vals[13];
    bool any_free = false;
    for (int a = 0; a < 3; ++a) {
        const bool fr = !dirichlet_velocity(tag[nd[a]]);
        lam_u[a] = fr ? lu[nd[a]] : 0.0;
        lam_v[a] = fr ? lv[nd[a]] : 0.0;
        any_free = any_free || fr;
        uh[a] = a1 * un[nd[a]] + a2 * um[nd[a]];
        vh[a] = a1 * vn[nd[a]] + a2 * vm[nd[a]];
        pe[a] = p[nd[a]];
        xs_u[a] = us[nd[a]];
        xs_v[a] = vs[nd[a]];
        vals[a] = x[nd[a]];
        vals[3 + a] = y[nd[a]];
        vals[6 + a] = c1 * un[nd[a]] + c2 * um[nd[a]];
        vals[9 + a] = c1 * vn[nd[a]] + c2 * vm[nd[a]];
    }
    vals[12] = nu;
    double grad[13] = {0}, lin[9] = {0};
    const bool ok = any_free && element_gradient(13, geom ? 0 : 6, vals, grad, [&](const D* in, bool value_pass, bool& good) {
        GeoD g;
        if (!geometry_d(in, in + 3, g)) {
            good = false;
            return dconst<ND>(0.0);
        }
        const D* WU = in + 6;
        const D* WV = in + 9;
        const D NU = in[12];
        const D area = g.area;
        D wu_bar = dconst<ND>(0.0), wv_bar = dconst<ND>(0.0), div_w = dconst<ND>(0.0);
        for (int k = 0; k < 3; ++k) {
            wu_bar = wu_bar + WU[k] / 3.0;
            wv_bar = wv_bar + WV[k] / 3.0;
            div_w = div_w + WU[k] * g.dNdx[k] + WV[k] * g.dNdy[k];
        }
        const D vel = dsqrt(wu_bar * wu_bar + wv_bar * wv_bar) + 1.0e-12;
        const D hK = dsqrt(2.0 * area);
        const double tt = 2.0 / tau_time;
        const D tc = 2.0 * vel / hK;
        const D td = 4.0 * NU / (hK * hK);
        const D tau = 1.0 / dsqrt(tt * tt + tc * tc + 9.0 * (td * td));
        D dpdx = dconst<ND>(0.0), dpdy = dconst<ND>(0.0);
        for (int k = 0; k < 3; ++k) {
            dpdx = dpdx + pe[k] * g.dNdx[k];
            dpdy = dpdy + pe[k] * g.dNdy[k];
        }
        dpdx = dpdx / rho;
        dpdy = dpdy / rho;
        D ug[3];
        for (int i = 0; i < 3; ++i) ug[i] = wu_bar * g.dNdx[i] + wv_bar * g.dNdy[i];
        D phi = dconst<ND>(0.0);
        for (int i = 0; i < 3; ++i) {
            if (lam_u[i] == 0.0 && lam_v[i] == 0.0) continue;
            const D Ms = (area / 3.0) * (tau * ug[i]);
            const D tw = area / 3.0 + area * (tau * ug[i]);
            D ru = dconst<ND>(0.0), rv = dconst<ND>(0.0);
            for (int j = 0; j < 3; ++j) {
                const D M = (i == j) ? area / 6.0 : area / 12.0;
                const D K = area * (g.dNdx[i] * g.dNdx[j] + g.dNdy[i] * g.dNdy[j]);
                const D Aij = a0 * (M + Ms) + K * NU + (area / 3.0) * ug[j] + (skew * M) * div_w +
                              area * (tau * (ug[i] * ug[j]));
                const D mj = M + Ms;
                ru = ru + uh[j] * mj - xs_u[j] * Aij;
                rv = rv + vh[j] * mj - xs_v[j] * Aij;
                if (value_pass) {
                    lin[j] += lam_u[i] * mj.v;
                    lin[3 + j] += lam_v[i] * mj.v;
                }
            }
            ru = ru - tw * dpdx;
            rv = rv - tw * dpdy;
            phi = phi + lam_u[i] * ru + lam_v[i] * rv;
            if (value_pass)
                for (int k = 0; k < 3; ++k)
                    lin[6 + k] -= tw.v * (lam_u[i] * g.dNdx[k].v + lam_v[i] * g.dNdy[k].v) / rho;
        }
        return phi;
    });
    if (!ok) {
        for (int f = 0; f < 7; ++f)
            for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = 0.0;
        nu_part[e] = 0.0;
        return;
    }
    for (int k = 0; k < 3; ++k) {
        contrib[0 * stride + 3 * e + k] = c1 * grad[6 + k] + a1 * lin[k];
        contrib[1 * stride + 3 * e + k] = c1 * grad[9 + k] + a1 * lin[3 + k];
        contrib[2 * stride + 3 * e + k] = c2 * grad[6 + k] + a2 * lin[k];
        contrib[3 * stride + 3 * e + k] = c2 * grad[9 + k] + a2 * lin[3 + k];
        contrib[4 * stride + 3 * e + k] = lin[6 + k];
        contrib[5 * stride + 3 * e + k] = geom ? grad[k] : 0.0;
        contrib[6 * stride + 3 * e + k] = geom ? grad[3 + k] : 0.0;
    }
    nu_part[e] = grad[12];
}

// ---- pressure: Phi_K = element part of lambda~^T (rhs - Op dp), Op = K + S(s),
// rhs = -(rho/dt) D u* - S(s) p, S(s) = sum_K s_K K_K - B(s)^T M^-1 B(s) (the OSS stabilisation),
// lambda~ zero on outlet rows, y = p + dp. The M^-1 coupling is linearised around the nodal fields
// hy = M^-1 B y and hl = M^-1 B lambda~:
//   Phi_K = -(rho/dt)|K|/3 sum(lambda~) div u* - s_K |K| gl.gy - |K| gl.gdp
//           + sum_k [w_K (gl.hy_k + gy.hl_k) - (hl_k.hy_k) |K|/3],  w_K = sqrt(s_K) |K|/3,
// with s_K = s(mean u*, nu, |K|) as k_oss_weights (fem_kernels.cuh:697-719).
// inputs: X[3], Y[3], US[3], VS[3], nu; y enters linearly (p^n derivative from the value pass).
// contrib fields: 0 u*, 1 v*, 2 p, 3 x, 4 y; nu_part[e]
__global__ void k_adj_pressure_elems(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ us,
                                     const double* __restrict__ vs, const double* __restrict__ lam,
                                     const double* __restrict__ yv, const double* __restrict__ dp,
                                     const double* __restrict__ hyx, const double* __restrict__ hyy,
                                     const double* __restrict__ hlx, const double* __restrict__ hly, double nu,
                                     double rho, double dt, double stab, int geom, double* __restrict__ contrib,
                                     double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double l[3], yk[3], dpk[3], hy[3][2], hl[3][2], vals[13];
    for (int k = 0; k < 3; ++k) {
        l[k] = lam[nd[k]];
        yk[k] = yv[nd[k]];
        dpk[k] = dp[nd[k]];
        hy[k][0] = stab > 0.0 ? hyx[nd[k]] : 0.0;
        hy[k][1] = stab > 0.0 ? hyy[nd[k]] : 0.0;
        hl[k][0] = stab > 0.0 ? hlx[nd[k]] : 0.0;
        hl[k][1] = stab > 0.0 ? hly[nd[k]] : 0.0;
        vals[k] = x[nd[k]];
        vals[3 + k] = y[nd[k]];
        vals[6 + k] = us[nd[k]];
        vals[9 + k] = vs[nd[k]];
    }
    vals[12] = nu;
    double grad[13] = {0}, lin[3] = {0};
    const bool ok = element_gradient(13, geom ? 0 : 6, vals, grad, [&](const D* in, bool value_pass, bool& good) {
        GeoD g;
        if (!geometry_d(in, in + 3, g)) {
            good = false;
            return dconst<ND>(0.0);
        }
        const D* US = in + 6;
        const D* VS = in + 9;
        const D NU = in[12];
        const D area = g.area;
        const double Lam = l[0] + l[1] + l[2];
        D div = dconst<ND>(0.0), glx = dconst<ND>(0.0), gly = dconst<ND>(0.0), gyx = dconst<ND>(0.0),
          gyy = dconst<ND>(0.0), gdx = dconst<ND>(0.0), gdy = dconst<ND>(0.0);
        for (int k = 0; k < 3; ++k) {
            div = div + US[k] * g.dNdx[k] + VS[k] * g.dNdy[k];
            glx = glx + l[k] * g.dNdx[k];
            gly = gly + l[k] * g.dNdy[k];
            gyx = gyx + yk[k] * g.dNdx[k];
            gyy = gyy + yk[k] * g.dNdy[k];
            gdx = gdx + dpk[k] * g.dNdx[k];
            gdy = gdy + dpk[k] * g.dNdy[k];
        }
        D phi = (-(rho / dt) * Lam) * ((area / 3.0) * div) - area * (glx * gdx + gly * gdy);
        if (stab > 0.0) {
            // k_oss_weights, fem_kernels.cuh:697-719
            const D h = dsqrt(2.0 * area);
            const D ub = (US[0] + US[1] + US[2]) / 3.0;
            const D vb = (VS[0] + VS[1] + VS[2]) / 3.0;
            const D q = 4.0 * NU / (h * h) + 2.0 * dsqrt(ub * ub + vb * vb) / h;
            const D tau = stab / (rho * q + 1.0e-300);
            const D s = (rho / dt) * tau;
            const D w = dsqrt(s) * area / 3.0;
            phi = phi - s * area * (glx * gyx + gly * gyy);
            double hysx = 0.0, hysy = 0.0;
            for (int k = 0; k < 3; ++k) {
                phi = phi + w * (glx * hy[k][0] + gly * hy[k][1]) + w * (gyx * hl[k][0] + gyy * hl[k][1]) -
                      (hl[k][0] * hy[k][0] + hl[k][1] * hy[k][1]) * (area / 3.0);
                hysx += hl[k][0];
                hysy += hl[k][1];
            }
            if (value_pass)
                for (int k = 0; k < 3; ++k)
                    lin[k] = -s.v * area.v * (g.dNdx[k].v * glx.v + g.dNdy[k].v * gly.v) +
                             w.v * (g.dNdx[k].v * hysx + g.dNdy[k].v * hysy);
        }
        return phi;
    });
    if (!ok) {
        for (int f = 0; f < 5; ++f)
            for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = 0.0;
        nu_part[e] = 0.0;
        return;
    }
    for (int k = 0; k < 3; ++k) {
        contrib[0 * stride + 3 * e + k] = grad[6 + k];
        contrib[1 * stride + 3 * e + k] = grad[9 + k];
        contrib[2 * stride + 3 * e + k] = lin[k];
        contrib[3 * stride + 3 * e + k] = geom ? grad[k] : 0.0;
        contrib[4 * stride + 3 * e + k] = geom ? grad[3 + k] : 0.0;
    }
    nu_part[e] = grad[12];
}

// ---- correction and rotational update, coordinate sensitivities --------------------------------
// u1 = u* - (dt/rho) G dp / m (corrected rows), p1 = p + dp + nu dt K dp / m: with z = mask u1_bar
// / m, q = p1_bar / m and c_i = [-(dt/rho) z_i.(G dp)_i + nu dt q_i (K dp)_i] / m_i held fixed,
//   Phi_K = sum_i [-(dt/rho) z_i.grad dp_K |K|/3 + nu dt q_i |K| grad N_i.grad dp_K - c_i |K|/3]
// inputs X[3], Y[3]; contrib fields: 0 x, 1 y
__global__ void k_adj_corr_geom_elems(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                                      const double* __restrict__ y, const double* __restrict__ dp,
                                      const double* __restrict__ zx, const double* __restrict__ zy,
                                      const double* __restrict__ q, const double* __restrict__ c, double nu,
                                      double dt, double rho, double* __restrict__ contrib) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double dpk[3], zxk[3], zyk[3], qk[3], ck[3], vals[6];
    for (int k = 0; k < 3; ++k) {
        dpk[k] = dp[nd[k]];
        zxk[k] = zx[nd[k]];
        zyk[k] = zy[nd[k]];
        qk[k] = q[nd[k]];
        ck[k] = c[nd[k]];
        vals[k] = x[nd[k]];
        vals[3 + k] = y[nd[k]];
    }
    double grad[6] = {0};
    const bool ok = element_gradient(6, 0, vals, grad, [&](const D* in, bool, bool& good) {
        GeoD g;
        if (!geometry_d(in, in + 3, g)) {
            good = false;
            return dconst<ND>(0.0);
        }
        D gdx = dconst<ND>(0.0), gdy = dconst<ND>(0.0);
        for (int k = 0; k < 3; ++k) {
            gdx = gdx + dpk[k] * g.dNdx[k];
            gdy = gdy + dpk[k] * g.dNdy[k];
        }
        D phi = dconst<ND>(0.0);
        for (int i = 0; i < 3; ++i)
            phi = phi - (dt / rho) * (zxk[i] * gdx + zyk[i] * gdy) * (g.area / 3.0) +
                  (nu * dt * qk[i]) * g.area * (g.dNdx[i] * gdx + g.dNdy[i] * gdy) - ck[i] * (g.area / 3.0);
        return phi;
    });
    for (int k = 0; k < 3; ++k) {
        contrib[0 * stride + 3 * e + k] = ok ? grad[k] : 0.0;
        contrib[1 * stride + 3 * e + k] = ok ? grad[3 + k] : 0.0;
    }
}

// c_i of the correction functional above; Kdp = -(rho/dt) div_num, g = G dp (nodal)
__global__ void k_adj_corr_c(int n, const double* __restrict__ mass, const double* __restrict__ zx,
                             const double* __restrict__ zy, const double* __restrict__ q,
                             const double* __restrict__ gx, const double* __restrict__ gy,
                             const double* __restrict__ div_num, double nu, double dt, double rho,
                             double* __restrict__ c) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = fmax(mass[i], 1.0e-14);
    const double kdp = -(rho / dt) * div_num[i];
    c[i] = (-(dt / rho) * (zx[i] * gx[i] + zy[i] * gy[i]) + nu * dt * q[i] * kdp) / m;
}

// ---- volume forces: gradient of wx F_x + wy F_y of one element (fem_kernels.cuh:1046-1123) -----
// inputs (28): X[3], Y[3], then u, v, u_old, v_old, u_prev, v_prev, p at the three corners, nu
// contrib fields: 0 u, 1 v, 2 u_old, 3 v_old, 4 u_prev, 5 v_prev, 6 p, 7 x, 8 y; nu_part[e].
// Elements that do not touch the cylinder write zeros.
__global__ void k_adj_forces_elems(int ne, const int* __restrict__ ev, const int* __restrict__ tag,
                                   const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ uo, const double* __restrict__ vo,
                                   const double* __restrict__ up, const double* __restrict__ vp,
                                   const double* __restrict__ p, double b0, double b1, double b2, double nu,
                                   double rho, double wx, double wy, int geom, double* __restrict__ contrib,
                                   double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double phi_c[3];
    bool touches = false;
    for (int i = 0; i < 3; ++i) {
        phi_c[i] = (tag[nd[i]] == BC_CYLINDER) ? 1.0 : 0.0;
        touches = touches || phi_c[i] > 0.0;
    }
    double vals[28], grad[28] = {0};
    const double* src[7] = {u, v, uo, vo, up, vp, p};
    for (int k = 0; k < 3; ++k) {
        vals[k] = x[nd[k]];
        vals[3 + k] = y[nd[k]];
    }
    for (int f = 0; f < 7; ++f)
        for (int k = 0; k < 3; ++k) vals[6 + 3 * f + k] = src[f][nd[k]];
    vals[27] = nu;
    const bool ok = touches && element_gradient(28, geom ? 0 : 6, vals, grad, [&](const D* in, bool, bool& good) {
        GeoD g;
        if (!geometry_d(in, in + 3, g)) {
            good = false;
            return dconst<ND>(0.0);
        }
        const D* ue = in + 6;
        const D* ve = in + 9;
        const D* uoe = in + 12;
        const D* voe = in + 15;
        const D* upe = in + 18;
        const D* vpe = in + 21;
        const D* pp = in + 24;
        const D NU = in[27];
        D ute[3], vte[3];
        D dudx = dconst<ND>(0.0), dudy = dconst<ND>(0.0), dvdx = dconst<ND>(0.0), dvdy = dconst<ND>(0.0);
        D dphidx = dconst<ND>(0.0), dphidy = dconst<ND>(0.0), pbar = dconst<ND>(0.0);
        for (int i = 0; i < 3; ++i) {
            ute[i] = b0 * ue[i] + b1 * uoe[i] + b2 * upe[i];
            vte[i] = b0 * ve[i] + b1 * voe[i] + b2 * vpe[i];
            dudx = dudx + ue[i] * g.dNdx[i];
            dudy = dudy + ue[i] * g.dNdy[i];
            dvdx = dvdx + ve[i] * g.dNdx[i];
            dvdy = dvdy + ve[i] * g.dNdy[i];
            dphidx = dphidx + phi_c[i] * g.dNdx[i];
            dphidy = dphidy + phi_c[i] * g.dNdy[i];
            pbar = pbar + pp[i] / 3.0;
        }
        D fx = dconst<ND>(0.0), fy = dconst<ND>(0.0);
        for (int i = 0; i < 3; ++i) {
            const int j = (i + 1) % 3;
            const double pm = 0.5 * (phi_c[i] + phi_c[j]);
            if (pm == 0.0) continue;
            const D um = 0.5 * (ue[i] + ue[j]), vm = 0.5 * (ve[i] + ve[j]);
            const D cx = um * dudx + vm * dudy, cy = um * dvdx + vm * dvdy;
            fx = fx + pm * (0.5 * (ute[i] + ute[j]) + cx);
            fy = fy + pm * (0.5 * (vte[i] + vte[j]) + cy);
        }
        fx = (g.area / 3.0) * fx;
        fy = (g.area / 3.0) * fy;
        fx = fx + g.area * (NU * (dudx * dphidx + dudy * dphidy) - (pbar / rho) * dphidx);
        fy = fy + g.area * (NU * (dvdx * dphidx + dvdy * dphidy) - (pbar / rho) * dphidy);
        return wx * fx + wy * fy;
    });
    for (int f = 0; f < 7; ++f)
        for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = ok ? grad[6 + 3 * f + k] : 0.0;
    for (int k = 0; k < 3; ++k) {
        contrib[7 * stride + 3 * e + k] = ok && geom ? grad[k] : 0.0;
        contrib[8 * stride + 3 * e + k] = ok && geom ? grad[3 + k] : 0.0;
    }
    nu_part[e] = ok ? grad[27] : 0.0;
}

// out_f[i] += sum over the elements of node i, in element order, of contrib_f[3 e + corner]
struct OutPtrs {
    double* p[9];
};
__global__ void k_adj_gather(int n, int ne, int nf, const int* __restrict__ inc_ptr,
                             const int* __restrict__ inc_elem, const double* __restrict__ contrib, OutPtrs out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t stride = 3 * (size_t)ne;
    for (int f = 0; f < nf; ++f) {
        double acc = 0.0;
        for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) acc += contrib[f * stride + inc_elem[s]];
        out.p[f][i] += acc;
    }
}

// theta_bar[slot] += sum of part[0..count) (one workgroup, fixed order)
__global__ void k_adj_accumulate(int count, const double* __restrict__ part, double* __restrict__ theta_bar,
                                 int slot) {
    __shared__ double lds[BLOCK / WAVE];
    const double s = sum_partials(part, count, 1, lds);
    if (threadIdx.x == 0) theta_bar[slot] += s;
}

// Inflow-peak sensitivity: the inlet rows' right-hand side is u_max 4 y (H - y) / H^2
// (fem_kernels.cuh:431-453), so d/du_max = sum over inlet rows of lambda_u 4 y (H - y) / H^2.
__global__ void k_adj_inlet_part(int n, const int* __restrict__ tag, const double* __restrict__ y,
                                 const double* __restrict__ lu, double H, double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    part[i] = tag[i] == BC_INLET ? lu[i] * 4.0 * y[i] * (H - y[i]) / (H * H) : 0.0;
}

// ---- correction and rotational pressure update ---------------------------------------------------
// p1 = p + dp + nu dt M^-1 K dp;  u1 = u* - (dt/rho) M^-1 Gx dp on interior and outlet rows.
// Nodal step: q = p1_bar / m, z = mask u1_bar / m, and the nu partials p1_bar (-rho div_num / m)
// with div_num = -(dt / rho) K dp.
__global__ void k_adj_corr_prep(int n, const int* __restrict__ tag, const double* __restrict__ mass,
                                const double* __restrict__ p1b, const double* __restrict__ u1b,
                                const double* __restrict__ v1b, const double* __restrict__ div_num, double rho,
                                double* __restrict__ q, double* __restrict__ zx, double* __restrict__ zy,
                                double* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = fmax(mass[i], 1.0e-14);
    const int t = tag[i];
    const bool corrected = t == BC_INTERIOR || t == BC_OUTLET;
    q[i] = p1b[i] / m;
    zx[i] = corrected ? u1b[i] / m : 0.0;
    zy[i] = corrected ? v1b[i] / m : 0.0;
    part[i] = p1b[i] * (-rho * div_num[i] / m);
}

// dp_bar_k = p1_bar_k + sum_K [nu dt |K| grad N_k . grad q_K - (dt/rho) |K|/3 (dN_k/dx Zx_K + dN_k/dy Zy_K)]
// with Z_K the corner sums of z; zero on outlet rows (dp is fixed to 0 there).
__global__ void k_adj_dp_rows(int n, const int* __restrict__ ev, const double* __restrict__ x,
                              const double* __restrict__ y, const int* __restrict__ tag,
                              const int* __restrict__ inc_ptr, const int* __restrict__ inc_elem,
                              const double* __restrict__ p1b, const double* __restrict__ q,
                              const double* __restrict__ zx, const double* __restrict__ zy, double nu, double dt,
                              double rho, double* __restrict__ dpb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (tag[i] == BC_OUTLET) {
        dpb[i] = 0.0;
        return;
    }
    double acc = p1b[i];
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int ea = inc_elem[s];
        const int e = ea / 3, a = ea - 3 * e;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd, g)) continue;
        double gqx = 0.0, gqy = 0.0, Zx = 0.0, Zy = 0.0;
        for (int k = 0; k < 3; ++k) {
            gqx += q[nd[k]] * g.dNdx[k];
            gqy += q[nd[k]] * g.dNdy[k];
            Zx += zx[nd[k]];
            Zy += zy[nd[k]];
        }
        acc += nu * dt * g.area * (g.dNdx[a] * gqx + g.dNdy[a] * gqy) -
               (dt / rho) * (g.area / 3.0) * (g.dNdx[a] * Zx + g.dNdy[a] * Zy);
    }
    dpb[i] = acc;
}

// y = p + dp; lam~ = lam with outlet rows zeroed
__global__ void k_adj_pressure_prep(int n, const int* __restrict__ tag, const double* __restrict__ p,
                                    const double* __restrict__ dp, const double* __restrict__ lam,
                                    double* __restrict__ yv, double* __restrict__ lt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    yv[i] = p[i] + dp[i];
    lt[i] = tag[i] == BC_OUTLET ? 0.0 : lam[i];
}

// h = g / max(mass, 1e-14) for both components, in place
__global__ void k_adj_div_mass(int n, const double* __restrict__ mass, double* __restrict__ gx,
                               double* __restrict__ gy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = fmax(mass[i], 1.0e-14);
    gx[i] /= m;
    gy[i] /= m;
}

// B-weighted gradient of f (fem_kernels.cuh:726-757, weights sqrt(s_K) |K| / 3; |K| / 3 when s_el
// is NULL, the correction's G f), row gather
__global__ void k_adj_weighted_gradient(int n, const int* __restrict__ ev, const double* __restrict__ x,
                                        const double* __restrict__ y, const int* __restrict__ inc_ptr,
                                        const int* __restrict__ inc_elem, const double* __restrict__ f,
                                        const double* __restrict__ s_el, double* __restrict__ gx,
                                        double* __restrict__ gy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0, ay = 0.0;
    for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
        const int e = inc_elem[s] / 3;
        const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
        Geo g;
        if (!geometry(x, y, nd, g)) continue;
        const double fx = f[nd[0]] * g.dNdx[0] + f[nd[1]] * g.dNdx[1] + f[nd[2]] * g.dNdx[2];
        const double fy = f[nd[0]] * g.dNdy[0] + f[nd[1]] * g.dNdy[1] + f[nd[2]] * g.dNdy[2];
        const double w = (s_el ? sqrt(s_el[e]) : 1.0) * g.area / 3.0;
        ax += w * fx;
        ay += w * fy;
    }
    gx[i] = ax;
    gy[i] = ay;
}

__global__ void k_gather_values(int nnz, const int* __restrict__ perm, const double* __restrict__ in,
                                double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) out[k] = in[perm[k]];
}

__global__ void k_scale(int n, double a, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}

bool mesh_ok(const cfa_mesh_t* m) {
    return m && m->n_nodes >= 0 && m->n_elems >= 0 && m->elem_v && m->x && m->y && m->bc_tag && m->inc_ptr &&
           m->inc_elem;
}

constexpr int ADJ_OUT_BLOCKS = 1;

void gather(const cfa_mesh_t* m, int nf, const double* contrib, double* const* outs, hipStream_t st) {
    OutPtrs o{};
    for (int f = 0; f < nf; ++f) o.p[f] = outs[f];
    hipLaunchKernelGGL(k_adj_gather, dim3(blocks_for(m->n_nodes)), dim3(BLOCK), 0, st, m->n_nodes, m->n_elems, nf,
                       m->inc_ptr, m->inc_elem, contrib, o);
}

void accumulate(int count, const double* part, double* theta_bar, int slot, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_accumulate, dim3(ADJ_OUT_BLOCKS), dim3(BLOCK), 0, st, count, part, theta_bar, slot);
}

}  // namespace

extern "C" {

size_t cfa_adj_workspace_doubles(int n, int ne) {
    if (n < 0 || ne < 0) return 0;
    return 28 * (size_t)ne + 9 * (size_t)n + 64;
}

namespace {
// workspace layout: per-(element, corner) fields [27 ne], partials [ne + n], nodal scratch [8 n]
struct Ws {
    double* contrib;
    double* part;
    double* node;
};
Ws split(double* ws, int n, int ne) {
    return {ws, ws + 27 * (size_t)ne, ws + 28 * (size_t)ne + n};
}
}  // namespace

int cfa_adj_momentum(const cfa_mesh_t* m, const double* u_n, const double* v_n, const double* u_prev,
                     const double* v_prev, const double* p, const double* u_star, const double* v_star,
                     const double* lam_u, const double* lam_v, double nu, double rho, double a0, double a1,
                     double a2, double c1, double c2, double tau_time, double skew, double h_channel,
                     double* ws, double* u_n_bar, double* v_n_bar, double* u_prev_bar, double* v_prev_bar,
                     double* p_bar, double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u_n || !v_n || !u_prev || !v_prev || !p || !u_star || !v_star || !lam_u || !lam_v || !ws ||
        !u_n_bar || !v_n_bar || !u_prev_bar || !v_prev_bar || !p_bar || !theta_bar || (!x_bar != !y_bar))
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    const Ws w = split(ws, n, ne);
    const int geom = x_bar ? 1 : 0;
    hipLaunchKernelGGL(k_adj_momentum_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->x, m->y,
                       m->bc_tag, u_n, v_n, u_prev, v_prev, p, u_star, v_star, lam_u, lam_v, nu, rho, a0, a1, a2,
                       c1, c2, tau_time, skew, geom, w.contrib, w.part);
    double* outs[7] = {u_n_bar, v_n_bar, u_prev_bar, v_prev_bar, p_bar, x_bar, y_bar};
    gather(m, geom ? 7 : 5, w.contrib, outs, st);
    accumulate(ne, w.part, theta_bar, 0, st);
    hipLaunchKernelGGL(k_adj_inlet_part, dim3(blocks_for(n)), dim3(BLOCK), 0, st, n, m->bc_tag, m->y, lam_u,
                       h_channel, w.part);
    accumulate(n, w.part, theta_bar, 1, st);
    return launch_status();
}

int cfa_adj_pressure(const cfa_mesh_t* m, const double* elem_s, const double* u_star, const double* v_star,
                     const double* p, const double* dp, const double* lam, const double* mass, double nu, double rho,
                     double dt, double stab, double* ws, double* u_star_bar, double* v_star_bar, double* p_bar,
                     double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !elem_s || !u_star || !v_star || !p || !dp || !lam || !mass || !ws || !u_star_bar ||
        !v_star_bar || !p_bar || !theta_bar || (!x_bar != !y_bar))
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    const Ws w = split(ws, n, ne);
    double* yv = w.node;
    double* lt = yv + n;
    double* hyx = lt + n;
    double* hyy = hyx + n;
    double* hlx = hyy + n;
    double* hly = hlx + n;
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k_adj_pressure_prep, dim3(nb), dim3(BLOCK), 0, st, n, m->bc_tag, p, dp, lam, yv, lt);
    if (stab > 0.0) {
        hipLaunchKernelGGL(k_adj_weighted_gradient, dim3(nb), dim3(BLOCK), 0, st, n, m->elem_v, m->x, m->y,
                           m->inc_ptr, m->inc_elem, yv, elem_s, hyx, hyy);
        hipLaunchKernelGGL(k_adj_weighted_gradient, dim3(nb), dim3(BLOCK), 0, st, n, m->elem_v, m->x, m->y,
                           m->inc_ptr, m->inc_elem, lt, elem_s, hlx, hly);
        hipLaunchKernelGGL(k_adj_div_mass, dim3(nb), dim3(BLOCK), 0, st, n, mass, hyx, hyy);
        hipLaunchKernelGGL(k_adj_div_mass, dim3(nb), dim3(BLOCK), 0, st, n, mass, hlx, hly);
    }
    const int geom = x_bar ? 1 : 0;
    hipLaunchKernelGGL(k_adj_pressure_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->x, m->y,
                       u_star, v_star, lt, yv, dp, hyx, hyy, hlx, hly, nu, rho, dt, stab, geom, w.contrib, w.part);
    double* outs[5] = {u_star_bar, v_star_bar, p_bar, x_bar, y_bar};
    gather(m, geom ? 5 : 3, w.contrib, outs, st);
    accumulate(ne, w.part, theta_bar, 0, st);
    return launch_status();
}

int cfa_adj_correction(const cfa_mesh_t* m, double dt, double rho, double nu, const double* div_num,
                       const double* mass, const double* dp, const double* p1_bar, const double* u1_bar,
                       const double* v1_bar, double* ws, double* dp_bar, double* x_bar, double* y_bar,
                       double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !div_num || !mass || !dp || !p1_bar || !u1_bar || !v1_bar || !ws || !dp_bar || !theta_bar ||
        (!x_bar != !y_bar))
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    const Ws w = split(ws, n, ne);
    double* q = w.node;
    double* zx = q + n;
    double* zy = zx + n;
    double* gx = zy + n;
    double* gy = gx + n;
    double* c = gy + n;
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k_adj_corr_prep, dim3(nb), dim3(BLOCK), 0, st, n, m->bc_tag, mass, p1_bar, u1_bar, v1_bar,
                       div_num, rho, q, zx, zy, w.part);
    accumulate(n, w.part, theta_bar, 0, st);
    hipLaunchKernelGGL(k_adj_dp_rows, dim3(nb), dim3(BLOCK), 0, st, n, m->elem_v, m->x, m->y, m->bc_tag,
                       m->inc_ptr, m->inc_elem, p1_bar, q, zx, zy, nu, dt, rho, dp_bar);
    if (x_bar && ne > 0) {
        hipLaunchKernelGGL(k_adj_weighted_gradient, dim3(nb), dim3(BLOCK), 0, st, n, m->elem_v, m->x, m->y,
                           m->inc_ptr, m->inc_elem, dp, nullptr, gx, gy);
        hipLaunchKernelGGL(k_adj_corr_c, dim3(nb), dim3(BLOCK), 0, st, n, mass, zx, zy, q, gx, gy, div_num, nu, dt,
                           rho, c);
        hipLaunchKernelGGL(k_adj_corr_geom_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->x,
                           m->y, dp, zx, zy, q, c, nu, dt, rho, w.contrib);
        double* outs[2] = {x_bar, y_bar};
        gather(m, 2, w.contrib, outs, st);
    }
    return launch_status();
}

int cfa_adj_forces(const cfa_mesh_t* m, const double* u, const double* v, const double* u_old,
                   const double* v_old, const double* u_prev, const double* v_prev, const double* p, double b0,
                   double b1, double b2, double nu, double rho, double wx, double wy, double* ws, double* u_bar,
                   double* v_bar, double* u_old_bar, double* v_old_bar, double* u_prev_bar, double* v_prev_bar,
                   double* p_bar, double* x_bar, double* y_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u || !v || !u_old || !v_old || !u_prev || !v_prev || !p || !ws || !u_bar || !v_bar ||
        !u_old_bar || !v_old_bar || !u_prev_bar || !v_prev_bar || !p_bar || !theta_bar || (!x_bar != !y_bar))
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    const Ws w = split(ws, n, ne);
    const int geom = x_bar ? 1 : 0;
    hipLaunchKernelGGL(k_adj_forces_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->bc_tag, m->x,
                       m->y, u, v, u_old, v_old, u_prev, v_prev, p, b0, b1, b2, nu, rho, wx, wy, geom, w.contrib,
                       w.part);
    double* outs[9] = {u_bar, v_bar, u_old_bar, v_old_bar, u_prev_bar, v_prev_bar, p_bar, x_bar, y_bar};
    gather(m, geom ? 9 : 7, w.contrib, outs, st);
    accumulate(ne, w.part, theta_bar, 0, st);
    return launch_status();
}

int cfa_gather_values(int nnz, const int* perm, const double* in, double* out, cfa_stream_t stream) {
    if (nnz < 0 || (nnz > 0 && (!perm || !in || !out)) || in == out) return CFA_E_ARG;
    if (nnz == 0) return CFA_OK;
    hipLaunchKernelGGL(k_gather_values, dim3(blocks_for(nnz)), dim3(BLOCK), 0, as_stream(stream), nnz, perm, in, out);
    return launch_status();
}

int cfa_scale(int n, double a, double* x, cfa_stream_t stream) {
    if (n < 0 || (n > 0 && !x)) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(BLOCK), 0, as_stream(stream), n, a, x);
    return launch_status();
}

}  // extern "C"
