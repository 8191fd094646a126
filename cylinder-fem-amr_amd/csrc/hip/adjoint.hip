// Discrete adjoint of the BDF2 projection step for gfx950: vector-Jacobian products of the step's
// element assemblies and nodal updates. The forward step is the one of src/cylinder_fem_amr.cu:
// 277-538 (momentum fem_kernels.cuh:40-181 + :411-453, OSS weights :697-719, pressure :201-282
// + :460-531, OSS operator :763-885, rotational term :627-649, correction :536-622, volume forces
// :1046-1123); the reference has no adjoint.
//
// Each linear solve x = A(q)^{-1} b(q) of the step is differentiated through the scalar
//     Phi(q) = lambda^T (b(q) - A(q) x),   A^T lambda = x_bar,
// so q_bar = dPhi/dq with lambda and x held fixed. Phi is a sum of element terms; each element's
// term depends nonlinearly on a few element quantities only (the mean advecting velocity, its
// divergence, viscosity), which are carried as forward-mode dual numbers through a copy of the
// element formulas; its linear dependencies are differentiated by hand. Element contributions
// to nodes go to a per-(element, corner) scratch array and are summed per node over the node ->
// element table in element order (no atomics; the result does not depend on scheduling).
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
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N>& b) {
    Dual<N> r = b;
    r.v = a + b.v;
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
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r;
    r.v = sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}

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

__device__ __forceinline__ bool dirichlet_velocity(int t) {
    return t == BC_INLET || t == BC_WALL || t == BC_CYLINDER;
}

// ---- momentum: Phi_K = sum over free rows a of lambda_a . (b_a^K - sum_b A_ab^K x_b) ------------
// contrib fields (stride 3 ne): 0 u_n, 1 v_n, 2 u_{n-1}, 3 v_{n-1}, 4 p; nu_part[e]
__global__ void k_adj_momentum_elems(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                                     const double* __restrict__ y, const int* __restrict__ tag,
                                     const double* __restrict__ un, const double* __restrict__ vn,
                                     const double* __restrict__ um, const double* __restrict__ vm,
                                     const double* __restrict__ p, const double* __restrict__ us,
                                     const double* __restrict__ vs, const double* __restrict__ lu,
                                     const double* __restrict__ lv, double nu, double rho, double a0, double a1,
                                     double a2, double c1, double c2, double tau_time, double skew,
                                     double* __restrict__ contrib, double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    Geo g;
    double lam_u[3], lam_v[3];
    bool any_free = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool fr = !dirichlet_velocity(tag[nd[a]]);
        lam_u[a] = fr ? lu[nd[a]] : 0.0;
        lam_v[a] = fr ? lv[nd[a]] : 0.0;
        any_free = any_free || fr;
    }
    if (!any_free || !geometry(x, y, nd, g)) {
#pragma unroll
        for (int f = 0; f < 5; ++f)
#pragma unroll
            for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = 0.0;
        nu_part[e] = 0.0;
        return;
    }
    const double area = g.area;
    double wu_bar = 0.0, wv_bar = 0.0, div_w = 0.0, uh[3], vh[3], xs_u[3], xs_v[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double wu = c1 * un[nd[k]] + c2 * um[nd[k]], wv = c1 * vn[nd[k]] + c2 * vm[nd[k]];
        uh[k] = a1 * un[nd[k]] + a2 * um[nd[k]];
        vh[k] = a1 * vn[nd[k]] + a2 * vm[nd[k]];
        wu_bar += wu / 3.0;
        wv_bar += wv / 3.0;
        div_w += wu * g.dNdx[k] + wv * g.dNdy[k];
        xs_u[k] = us[nd[k]];
        xs_v[k] = vs[nd[k]];
    }
    const double dpdx = (p[nd[0]] * g.dNdx[0] + p[nd[1]] * g.dNdx[1] + p[nd[2]] * g.dNdx[2]) / rho;
    const double dpdy = (p[nd[0]] * g.dNdy[0] + p[nd[1]] * g.dNdy[1] + p[nd[2]] * g.dNdy[2]) / rho;
    // dual over X = (wu_bar, wv_bar, div_w, nu)
    using D = Dual<4>;
    D Wu = dconst<4>(wu_bar), Wv = dconst<4>(wv_bar), Dw = dconst<4>(div_w), Nu = dconst<4>(nu);
    Wu.d[0] = 1.0;
    Wv.d[1] = 1.0;
    Dw.d[2] = 1.0;
    Nu.d[3] = 1.0;
    const D vel = dsqrt(Wu * Wu + Wv * Wv) + dconst<4>(1.0e-12);
    const double hK = sqrt(2.0 * area);
    const double tt = 2.0 / tau_time;
    const D tc = (2.0 / hK) * vel;
    const D td = (4.0 / (hK * hK)) * Nu;
    const D tau = dconst<4>(1.0) / dsqrt(dconst<4>(tt * tt) + tc * tc + 9.0 * (td * td));
    D ug[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) ug[i] = g.dNdx[i] * Wu + g.dNdy[i] * Wv;
    D phi = dconst<4>(0.0);
    double dphi_duh[3] = {0.0, 0.0, 0.0}, dphi_dvh[3] = {0.0, 0.0, 0.0}, lam_tw_x = 0.0, lam_tw_y = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (lam_u[i] == 0.0 && lam_v[i] == 0.0) continue;
        const D Ms = (area / 3.0) * (tau * ug[i]);
        const D tw = (area / 3.0) + area * (tau * ug[i]);
        D ru = dconst<4>(0.0), rv = dconst<4>(0.0);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double M = (i == j) ? (area / 6.0) : (area / 12.0);
            const double K = area * (g.dNdx[i] * g.dNdx[j] + g.dNdy[i] * g.dNdy[j]);
            const D Aij = a0 * (M + Ms) + K * Nu + (area / 3.0) * ug[j] + (skew * M) * Dw +
                          area * (tau * (ug[i] * ug[j]));
            const D mj = M + Ms;
            ru = ru + uh[j] * mj - xs_u[j] * Aij;
            rv = rv + vh[j] * mj - xs_v[j] * Aij;
            dphi_duh[j] += lam_u[i] * mj.v;
            dphi_dvh[j] += lam_v[i] * mj.v;
        }
        ru = ru - dpdx * tw;
        rv = rv - dpdy * tw;
        phi = phi + lam_u[i] * ru + lam_v[i] * rv;
        lam_tw_x += tw.v * lam_u[i];
        lam_tw_y += tw.v * lam_v[i];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double gu = phi.d[0] / 3.0 + phi.d[2] * g.dNdx[k];
        const double gv = phi.d[1] / 3.0 + phi.d[2] * g.dNdy[k];
        contrib[0 * stride + 3 * e + k] = c1 * gu + a1 * dphi_duh[k];
        contrib[1 * stride + 3 * e + k] = c1 * gv + a1 * dphi_dvh[k];
        contrib[2 * stride + 3 * e + k] = c2 * gu + a2 * dphi_duh[k];
        contrib[3 * stride + 3 * e + k] = c2 * gv + a2 * dphi_dvh[k];
        contrib[4 * stride + 3 * e + k] = -(lam_tw_x * g.dNdx[k] + lam_tw_y * g.dNdy[k]) / rho;
    }
    nu_part[e] = phi.d[3];
}

// ---- pressure: Phi_K = lambda~^T (rhs^K - Op^K dp), Op = K + S(s), rhs = -(rho/dt) D u* - S(s) p
// with S(s) = sum_K s_K K_K - B(s)^T M^-1 B(s) (the OSS stabilisation), lambda~ zero on outlet
// rows. hy = M^-1 B (p + dp), hl = M^-1 B lambda~ (nodal, computed beforehand).
// contrib fields: 0 u*, 1 v*, 2 p; nu_part[e]
__global__ void k_adj_pressure_elems(int ne, const int* __restrict__ ev, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ s_el,
                                     const double* __restrict__ us, const double* __restrict__ vs,
                                     const double* __restrict__ lam, const double* __restrict__ yv,
                                     const double* __restrict__ hyx, const double* __restrict__ hyy,
                                     const double* __restrict__ hlx, const double* __restrict__ hly, double nu,
                                     double rho, double dt, double stab, double* __restrict__ contrib,
                                     double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    Geo g;
    if (!geometry(x, y, nd, g)) {
#pragma unroll
        for (int f = 0; f < 3; ++f)
#pragma unroll
            for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = 0.0;
        nu_part[e] = 0.0;
        return;
    }
    const double area = g.area;
    double glx = 0.0, gly = 0.0, gyx = 0.0, gyy = 0.0, Hyx = 0.0, Hyy = 0.0, Hlx = 0.0, Hly = 0.0, Lam = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double l = lam[nd[k]], yk = yv[nd[k]];
        glx += l * g.dNdx[k];
        gly += l * g.dNdy[k];
        gyx += yk * g.dNdx[k];
        gyy += yk * g.dNdy[k];
        Hyx += hyx[nd[k]];
        Hyy += hyy[nd[k]];
        Hlx += hlx[nd[k]];
        Hly += hly[nd[k]];
        Lam += l;
    }
    const double s = s_el[e];
    const double rs = sqrt(s);
    const double w = rs * area / 3.0;
    double s_bar = 0.0;
    if (stab > 0.0) {
        const double dw_ds = rs > 0.0 ? area / (6.0 * rs) : 0.0;
        s_bar = -(area * (glx * gyx + gly * gyy) - dw_ds * (glx * Hyx + gly * Hyy + gyx * Hlx + gyy * Hly));
    }
    // s_K(ubar, vbar, nu) as k_oss_weights (fem_kernels.cuh:697-719), dual over (ubar, vbar, nu)
    using D = Dual<3>;
    D Ub = dconst<3>((us[nd[0]] + us[nd[1]] + us[nd[2]]) / 3.0);
    D Vb = dconst<3>((vs[nd[0]] + vs[nd[1]] + vs[nd[2]]) / 3.0);
    D Nu = dconst<3>(nu);
    Ub.d[0] = 1.0;
    Vb.d[1] = 1.0;
    Nu.d[2] = 1.0;
    double ds[3] = {0.0, 0.0, 0.0};
    if (stab > 0.0) {
        const double h = sqrt(2.0 * area);
        const D q = (4.0 / (h * h)) * Nu + (2.0 / h) * dsqrt(Ub * Ub + Vb * Vb);
        const D tau = dconst<3>(stab) / (rho * q + dconst<3>(1.0e-300));
        const D sd = (rho / dt) * tau;
#pragma unroll
        for (int k = 0; k < 3; ++k) ds[k] = sd.d[k];
    }
    const double div_c = -(rho / dt) * (area / 3.0) * Lam;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        contrib[0 * stride + 3 * e + k] = s_bar * ds[0] / 3.0 + div_c * g.dNdx[k];
        contrib[1 * stride + 3 * e + k] = s_bar * ds[1] / 3.0 + div_c * g.dNdy[k];
        contrib[2 * stride + 3 * e + k] =
            stab > 0.0 ? -s * area * (g.dNdx[k] * glx + g.dNdy[k] * gly) + w * (g.dNdx[k] * Hlx + g.dNdy[k] * Hly)
                       : 0.0;
    }
    nu_part[e] = s_bar * ds[2];
}

// ---- volume forces: gradient of wx F_x + wy F_y of one element (fem_kernels.cuh:1046-1123) -----
// inputs (22): u, v, u_old, v_old, u_prev, v_prev, p at the three corners, then nu
template <typename T>
__device__ void element_force_t(const Geo& g, const double* phi, const T* in, double b0, double b1, double b2,
                                double rho, T& fx_out, T& fy_out) {
    const T* ue = in;
    const T* ve = in + 3;
    const T* uo = in + 6;
    const T* vo = in + 9;
    const T* up = in + 12;
    const T* vp = in + 15;
    const T* pp = in + 18;
    const T nu = in[21];
    T ute[3], vte[3];
    T dudx = dconst<4>(0.0), dudy = dconst<4>(0.0), dvdx = dconst<4>(0.0), dvdy = dconst<4>(0.0);
    T pbar = dconst<4>(0.0);
    double dphidx = 0.0, dphidy = 0.0;
    for (int i = 0; i < 3; ++i) {
        ute[i] = b0 * ue[i] + b1 * uo[i] + b2 * up[i];
        vte[i] = b0 * ve[i] + b1 * vo[i] + b2 * vp[i];
        dudx = dudx + g.dNdx[i] * ue[i];
        dudy = dudy + g.dNdy[i] * ue[i];
        dvdx = dvdx + g.dNdx[i] * ve[i];
        dvdy = dvdy + g.dNdy[i] * ve[i];
        dphidx += phi[i] * g.dNdx[i];
        dphidy += phi[i] * g.dNdy[i];
        pbar = pbar + (1.0 / 3.0) * pp[i];
    }
    T fx = dconst<4>(0.0), fy = dconst<4>(0.0);
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const double pm = 0.5 * (phi[i] + phi[j]);
        if (pm == 0.0) continue;
        const T um = 0.5 * (ue[i] + ue[j]), vm = 0.5 * (ve[i] + ve[j]);
        const T cx = um * dudx + vm * dudy, cy = um * dvdx + vm * dvdy;
        fx = fx + pm * (0.5 * (ute[i] + ute[j]) + cx);
        fy = fy + pm * (0.5 * (vte[i] + vte[j]) + cy);
    }
    fx = (g.area / 3.0) * fx;
    fy = (g.area / 3.0) * fy;
    fx = fx + g.area * (nu * (dphidx * dudx + dphidy * dudy) - (dphidx / rho) * pbar);
    fy = fy + g.area * (nu * (dphidx * dvdx + dphidy * dvdy) - (dphidy / rho) * pbar);
    fx_out = fx;
    fy_out = fy;
}

// contrib fields: 0 u, 1 v, 2 u_old, 3 v_old, 4 u_prev, 5 v_prev, 6 p; nu_part[e]. Elements that
// do not touch the cylinder write zeros.
__global__ void k_adj_forces_elems(int ne, const int* __restrict__ ev, const int* __restrict__ tag,
                                   const double* __restrict__ x, const double* __restrict__ y,
                                   const double* __restrict__ u, const double* __restrict__ v,
                                   const double* __restrict__ uo, const double* __restrict__ vo,
                                   const double* __restrict__ up, const double* __restrict__ vp,
                                   const double* __restrict__ p, double b0, double b1, double b2, double nu,
                                   double rho, double wx, double wy, double* __restrict__ contrib,
                                   double* __restrict__ nu_part) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const size_t stride = 3 * (size_t)ne;
    const int nd[3] = {ev[3 * e], ev[3 * e + 1], ev[3 * e + 2]};
    double phi[3];
    bool touches = false;
    for (int i = 0; i < 3; ++i) {
        phi[i] = (tag[nd[i]] == BC_CYLINDER) ? 1.0 : 0.0;
        touches = touches || phi[i] > 0.0;
    }
    Geo g;
    if (!touches || !geometry(x, y, nd, g)) {
        for (int f = 0; f < 7; ++f)
            for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = 0.0;
        nu_part[e] = 0.0;
        return;
    }
    const double* src[7] = {u, v, uo, vo, up, vp, p};
    double vals[22];
    for (int f = 0; f < 7; ++f)
        for (int k = 0; k < 3; ++k) vals[3 * f + k] = src[f][nd[k]];
    vals[21] = nu;
    double grad[22];
    using D = Dual<4>;
    for (int pass = 0; pass < 6; ++pass) {
        D in[22];
        for (int q = 0; q < 22; ++q) {
            in[q] = dconst<4>(vals[q]);
            const int slot = q - 4 * pass;
            if (slot >= 0 && slot < 4) in[q].d[slot] = 1.0;
        }
        D fx, fy;
        element_force_t<D>(g, phi, in, b0, b1, b2, rho, fx, fy);
        for (int slot = 0; slot < 4; ++slot) {
            const int q = 4 * pass + slot;
            if (q < 22) grad[q] = wx * fx.d[slot] + wy * fy.d[slot];
        }
    }
    for (int f = 0; f < 7; ++f)
        for (int k = 0; k < 3; ++k) contrib[f * stride + 3 * e + k] = grad[3 * f + k];
    nu_part[e] = grad[21];
}

// out_f[i] += sum over the elements of node i, in element order, of contrib_f[3 e + corner]
struct OutPtrs {
    double* p[7];
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

// B-weighted gradient of f (fem_kernels.cuh:726-757, weights sqrt(s_K) |K| / 3), row gather
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
        const double w = sqrt(s_el[e]) * g.area / 3.0;
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
    return 22 * (size_t)ne + 6 * (size_t)n + 64;
}

int cfa_adj_momentum(const cfa_mesh_t* m, const double* u_n, const double* v_n, const double* u_prev,
                     const double* v_prev, const double* p, const double* u_star, const double* v_star,
                     const double* lam_u, const double* lam_v, double nu, double rho, double a0, double a1,
                     double a2, double c1, double c2, double tau_time, double skew, double h_channel,
                     double* ws, double* u_n_bar, double* v_n_bar, double* u_prev_bar, double* v_prev_bar,
                     double* p_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u_n || !v_n || !u_prev || !v_prev || !p || !u_star || !v_star || !lam_u || !lam_v || !ws ||
        !u_n_bar || !v_n_bar || !u_prev_bar || !v_prev_bar || !p_bar || !theta_bar)
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    double* contrib = ws;
    double* part = ws + 21 * (size_t)ne;   // ne element partials, then n node partials fit in 6n + ne
    hipLaunchKernelGGL(k_adj_momentum_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->x, m->y,
                       m->bc_tag, u_n, v_n, u_prev, v_prev, p, u_star, v_star, lam_u, lam_v, nu, rho, a0, a1, a2,
                       c1, c2, tau_time, skew, contrib, part);
    double* outs[5] = {u_n_bar, v_n_bar, u_prev_bar, v_prev_bar, p_bar};
    gather(m, 5, contrib, outs, st);
    accumulate(ne, part, theta_bar, 0, st);
    hipLaunchKernelGGL(k_adj_inlet_part, dim3(blocks_for(n)), dim3(BLOCK), 0, st, n, m->bc_tag, m->y, lam_u,
                       h_channel, part);
    accumulate(n, part, theta_bar, 1, st);
    return launch_status();
}

int cfa_adj_pressure(const cfa_mesh_t* m, const double* elem_s, const double* u_star, const double* v_star,
                     const double* p, const double* dp, const double* lam, const double* mass, double nu, double rho,
                     double dt, double stab, double* ws, double* u_star_bar, double* v_star_bar, double* p_bar,
                     double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !elem_s || !u_star || !v_star || !p || !dp || !lam || !mass || !ws || !u_star_bar ||
        !v_star_bar || !p_bar || !theta_bar)
        return CFA_E_ARG;
    const int n = m->n_nodes, ne = m->n_elems;
    if (n == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    double* contrib = ws;
    double* part = ws + 21 * (size_t)ne;
    double* yv = ws + 22 * (size_t)ne;
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
    hipLaunchKernelGGL(k_adj_pressure_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->x, m->y,
                       elem_s, u_star, v_star, lt, yv, hyx, hyy, hlx, hly, nu, rho, dt, stab, contrib, part);
    double* outs[3] = {u_star_bar, v_star_bar, p_bar};
    gather(m, 3, contrib, outs, st);
    accumulate(ne, part, theta_bar, 0, st);
    return launch_status();
}

int cfa_adj_correction(const cfa_mesh_t* m, double dt, double rho, double nu, const double* div_num,
                       const double* mass, const double* p1_bar, const double* u1_bar, const double* v1_bar,
                       double* ws, double* dp_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !div_num || !mass || !p1_bar || !u1_bar || !v1_bar || !ws || !dp_bar || !theta_bar)
        return CFA_E_ARG;
    const int n = m->n_nodes;
    if (n == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    double* q = ws;
    double* zx = ws + n;
    double* zy = zx + n;
    double* part = zy + n;
    const int nb = blocks_for(n);
    hipLaunchKernelGGL(k_adj_corr_prep, dim3(nb), dim3(BLOCK), 0, st, n, m->bc_tag, mass, p1_bar, u1_bar, v1_bar,
                       div_num, rho, q, zx, zy, part);
    accumulate(n, part, theta_bar, 0, st);
    hipLaunchKernelGGL(k_adj_dp_rows, dim3(nb), dim3(BLOCK), 0, st, n, m->elem_v, m->x, m->y, m->bc_tag,
                       m->inc_ptr, m->inc_elem, p1_bar, q, zx, zy, nu, dt, rho, dp_bar);
    return launch_status();
}

int cfa_adj_forces(const cfa_mesh_t* m, const double* u, const double* v, const double* u_old,
                   const double* v_old, const double* u_prev, const double* v_prev, const double* p, double b0,
                   double b1, double b2, double nu, double rho, double wx, double wy, double* ws, double* u_bar,
                   double* v_bar, double* u_old_bar, double* v_old_bar, double* u_prev_bar, double* v_prev_bar,
                   double* p_bar, double* theta_bar, cfa_stream_t stream) {
    if (!mesh_ok(m) || !u || !v || !u_old || !v_old || !u_prev || !v_prev || !p || !ws || !u_bar || !v_bar ||
        !u_old_bar || !v_old_bar || !u_prev_bar || !v_prev_bar || !p_bar || !theta_bar)
        return CFA_E_ARG;
    const int ne = m->n_elems;
    if (m->n_nodes == 0 || ne == 0) return CFA_OK;
    const hipStream_t st = as_stream(stream);
    double* contrib = ws;
    double* part = ws + 21 * (size_t)ne;
    hipLaunchKernelGGL(k_adj_forces_elems, dim3(blocks_for(ne)), dim3(BLOCK), 0, st, ne, m->elem_v, m->bc_tag, m->x,
                       m->y, u, v, u_old, v_old, u_prev, v_prev, p, b0, b1, b2, nu, rho, wx, wy, contrib, part);
    double* outs[7] = {u_bar, v_bar, u_old_bar, v_old_bar, u_prev_bar, v_prev_bar, p_bar};
    gather(m, 7, contrib, outs, st);
    accumulate(ne, part, theta_bar, 0, st);
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
