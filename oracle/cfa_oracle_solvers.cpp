// ORACLE (test infrastructure only): Krylov solvers, multicolour ILU(0) and SA-AMG of
// src/cfa/linalg.cuh restated on the host. Stopping rules, breakdown guards and AMG constants are
// the reference's. Row loops run under OpenMP (each row is computed by one thread, so results do
// not depend on the thread count); dot products sum fixed 4096-entry chunks left to right and
// then the chunk sums left to right, the same order for any number of threads. The AMG set-up
// stays serial.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "cfa_oracle.hpp"

namespace oracle {

double dot(int n, const double* a, const double* b) {
    constexpr int CHUNK = 4096;
    const int nch = (n + CHUNK - 1) / CHUNK;
    Vec part(nch > 0 ? nch : 1, 0.0);
#pragma omp parallel for schedule(static) if (nch > 4)
    for (int c = 0; c < nch; ++c) {
        const int hi = std::min(n, (c + 1) * CHUNK);
        double s = 0.0;
        for (int i = c * CHUNK; i < hi; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[c];
    return s;
}

// src/cfa/linalg.cuh:8-26
void spmv(int n, const int* rp, const int* ci, const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[ci[k]];
        y[i] = s;
    }
}

// src/cfa/linalg.cuh:183-256: PCG, warm start, ||r|| < tol ||b||
SolveInfo pcg(int n, const int* rp, const int* ci, const double* val, const double* b, double* x,
              int max_iter, double tol, const Op* extra, const Op* prec) {
    SolveInfo info;
    Vec r(n), z(n), p(n), q(n), inv_diag(n);
    auto apply = [&](const double* in, double* out) {
        spmv(n, rp, ci, val, in, out);
        if (extra) (*extra)(in, out);
    };
    const double b_norm = std::sqrt(dot(n, b, b));
    if (b_norm < 1.0e-14) {
        std::fill(x, x + n, 0.0);
        info.iters = 0;
        return info;
    }
    for (int i = 0; i < n; ++i) {   // src/cfa/linalg.cuh:28-48
        double d = 1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { d = val[k]; break; }
        inv_diag[i] = (std::fabs(d) > 1.0e-14) ? (1.0 / d) : 1.0;
    }
    apply(x, q.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * q[i];
    double r_norm = std::sqrt(dot(n, r.data(), r.data()));
    if (r_norm < tol * b_norm) return info;
    auto apply_prec = [&]() {
        if (prec) (*prec)(r.data(), z.data());
        else
#pragma omp parallel for schedule(static) if (n > 16384)
            for (int i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
    };
    apply_prec();
    p = z;
    double rz_old = dot(n, r.data(), z.data());
    int iter = 0;
    bool done = false;
    for (; iter < max_iter; ++iter) {
        apply(p.data(), q.data());
        const double pq = dot(n, p.data(), q.data());
        if (std::fabs(pq) < 1.0e-30) break;
        const double alpha = rz_old / pq;
#pragma omp parallel for schedule(static) if (n > 16384)
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] += -alpha * q[i];
        }
        r_norm = std::sqrt(dot(n, r.data(), r.data()));
        info.rel = r_norm / b_norm;
        if (r_norm < tol * b_norm) {
            ++iter;
            done = true;
            break;
        }
        apply_prec();
        const double rz_new = dot(n, r.data(), z.data());
        const double beta = rz_new / (rz_old + 1.0e-30);
#pragma omp parallel for schedule(static) if (n > 16384)
        for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        rz_old = rz_new;
    }
    info.converged = done;
    info.iters = iter;
    return info;
}

// src/cfa/linalg.cuh:261-347: Jacobi-preconditioned BiCGSTAB, one right-hand side
SolveInfo bicgstab_jacobi(int n, const int* rp, const int* ci, const double* val, const double* b,
                          double* x, int max_iter, double tol) {
    SolveInfo info;
    Vec r(n), r0(n), p(n), v(n, 0.0), s(n), t(n), ph(n), sh(n), dinv(n);
    for (int i = 0; i < n; ++i) {
        double d = 1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) { d = val[k]; break; }
        dinv[i] = (std::fabs(d) > 1.0e-14) ? (1.0 / d) : 1.0;
    }
    spmv(n, rp, ci, val, x, v.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * v[i];
    r0 = r;
    p = r;
    std::fill(v.begin(), v.end(), 0.0);
    const double rn0 = std::sqrt(dot(n, r.data(), r.data()));
    if (rn0 < 1.0e-13) return info;
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    const double bd = 1.0e-30 * rn0 * rn0 + 1.0e-300;
    int iter = 0;
    bool done = false;
    for (; iter < max_iter; ++iter) {
        const double rho_new = dot(n, r0.data(), r.data());
        if (std::fabs(rho_new) < bd) break;
        if (iter > 0) {
            const double beta = (rho_new / rho_old) * (alpha / (omega + 1.0e-30));
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        for (int i = 0; i < n; ++i) ph[i] = dinv[i] * p[i];
        spmv(n, rp, ci, val, ph.data(), v.data());
        const double r0v = dot(n, r0.data(), v.data());
        if (std::fabs(r0v) < bd) break;
        alpha = rho_new / r0v;
        for (int i = 0; i < n; ++i) s[i] = r[i] + (-alpha) * v[i];
        const double s_norm = std::sqrt(dot(n, s.data(), s.data()));
        if (s_norm < tol * rn0 || s_norm < 1.0e-11) {
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            ++iter;
            done = true;
            break;
        }
        for (int i = 0; i < n; ++i) sh[i] = dinv[i] * s[i];
        spmv(n, rp, ci, val, sh.data(), t.data());
        const double tt = dot(n, t.data(), t.data());
        omega = (tt > bd) ? (dot(n, t.data(), s.data()) / tt) : 0.0;
        for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
        for (int i = 0; i < n; ++i) x[i] += omega * sh[i];
        for (int i = 0; i < n; ++i) r[i] = s[i] + (-omega) * t[i];
        const double r_norm = std::sqrt(dot(n, r.data(), r.data()));
        info.rel = r_norm / rn0;
        if (r_norm < tol * rn0 || r_norm < 1.0e-11) {
            ++iter;
            done = true;
            break;
        }
        rho_old = rho_new;
    }
    info.converged = done;
    info.iters = iter;
    return info;
}

// ---- multicolour ILU(0) (src/cfa/linalg.cuh:361-568) ----------------------------------------
void Ilu0::set_pattern(int n_, const int* rp_, const int* ci_) {
    n = n_;  rp = rp_;  ci = ci_;
    colour.assign(n, -1);
    diag.assign(n, -1);
    std::vector<int> mark;
    n_colours = 0;
    for (int i = 0; i < n; ++i) {
        if (rp[i + 1] - rp[i] > 32) {
            std::fprintf(stderr, "[ILU] row %d has %d entries (limit 32)\n", i, rp[i + 1] - rp[i]);
            std::exit(EXIT_FAILURE);
        }
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int c = colour[ci[a]];
            if (c >= 0) mark[c] = i;
            if (ci[a] == i) diag[i] = a;
        }
        int c = 0;
        while (c < n_colours && mark[c] == i) ++c;
        if (c == n_colours) { ++n_colours; mark.push_back(-1); }
        colour[i] = c;
    }
    colour_start.assign(n_colours + 1, 0);
    for (int i = 0; i < n; ++i) colour_start[colour[i] + 1]++;
    for (int c = 0; c < n_colours; ++c) colour_start[c + 1] += colour_start[c];
    rows.assign(n, 0);
    std::vector<int> fill(colour_start.begin(), colour_start.end() - 1);
    for (int i = 0; i < n; ++i) rows[fill[colour[i]]++] = i;
    lp.assign(n + 1, 0);
    up.assign(n + 1, 0);
    lc.clear(); uc.clear(); l_src.clear(); u_src.clear();
    d_src.assign(n, -1);
    for (int t = 0; t < n; ++t) {
        const int i = rows[t];
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int j = ci[a];
            if (j == i) d_src[t] = a;
            else if (colour[j] < colour[i]) { lc.push_back(j); l_src.push_back(a); }
            else { uc.push_back(j); u_src.push_back(a); }
        }
        lp[t + 1] = (int)lc.size();
        up[t + 1] = (int)uc.size();
    }
}

void Ilu0::factor(const double* val) {
    const int nnz = rp[n];
    lu.assign(val, val + nnz);
    for (int cc = 1; cc < n_colours; ++cc) {
        for (int t = colour_start[cc]; t < colour_start[cc + 1]; ++t) {
            const int i = rows[t];
            const int s = rp[i], len = rp[i + 1] - s;
            const int ci_i = colour[i];
            int cols[32], ccol[32];
            double a[32];
            for (int q = 0; q < len; ++q) { cols[q] = ci[s + q]; ccol[q] = colour[cols[q]]; a[q] = lu[s + q]; }
            for (int c = 0; c < ci_i; ++c)
                for (int q = 0; q < len; ++q) {
                    if (ccol[q] != c) continue;
                    const int k = cols[q];
                    const double lik = a[q] / lu[diag[k]];
                    a[q] = lik;
                    for (int m = rp[k]; m < rp[k + 1]; ++m) {
                        const int j = ci[m];
                        for (int b = 0; b < len; ++b)
                            if (cols[b] == j) { if (ccol[b] > c) a[b] -= lik * lu[m]; break; }
                    }
                }
            for (int q = 0; q < len; ++q) lu[s + q] = a[q];
        }
    }
    lv.resize(l_src.size());
    uv.resize(u_src.size());
    dinv.resize(n);
    for (size_t e = 0; e < l_src.size(); ++e) lv[e] = lu[l_src[e]];
    for (size_t e = 0; e < u_src.size(); ++e) uv[e] = lu[u_src[e]];
    for (int t = 0; t < n; ++t) dinv[t] = 1.0 / lu[d_src[t]];
}

void Ilu0::apply(const double* r, double* z, int nrhs) const {
    for (int comp = 0; comp < nrhs; ++comp) {
        const double* rc = r + (size_t)comp * n;
        double* zc = z + (size_t)comp * n;
        for (int c = 0; c < n_colours; ++c) {
            const bool last = (c == n_colours - 1);
#pragma omp parallel for schedule(static) if (n > 16384)
            for (int t = colour_start[c]; t < colour_start[c + 1]; ++t) {
                const int i = rows[t];
                double acc = rc[i];
                for (int e = lp[t]; e < lp[t + 1]; ++e) acc -= lv[e] * zc[lc[e]];
                zc[i] = last ? acc * dinv[t] : acc;
            }
        }
        for (int c = n_colours - 2; c >= 0; --c)
#pragma omp parallel for schedule(static) if (n > 16384)
            for (int t = colour_start[c]; t < colour_start[c + 1]; ++t) {
                const int i = rows[t];
                double acc = zc[i];
                for (int e = up[t]; e < up[t + 1]; ++e) acc -= uv[e] * zc[uc[e]];
                zc[i] = acc * dinv[t];
            }
    }
}

// src/cfa/linalg.cuh:660-754: two components sharing A, per-component coefficients and stopping
Bicg2Info bicgstab2(int n, const int* rp, const int* ci, const double* val, const double* bu,
                    const double* bv, double* xu, double* xv, int max_iter, double tol,
                    const Ilu0& ilu) {
    Bicg2Info info;
    const size_t N = 2 * (size_t)n;
    Vec x(N), b(N), r(N), r0(N), p(N), v(N), s(N), t(N), ph(N), sh(N);
    std::copy(xu, xu + n, x.begin());
    std::copy(xv, xv + n, x.begin() + n);
    std::copy(bu, bu + n, b.begin());
    std::copy(bv, bv + n, b.begin() + n);
    auto spmv2 = [&](const double* in, double* out) {
        spmv(n, rp, ci, val, in, out);
        spmv(n, rp, ci, val, in + n, out + n);
    };
    auto dot_c = [&](const Vec& a, const Vec& bb, int c) { return dot(n, a.data() + (size_t)c * n, bb.data() + (size_t)c * n); };
    spmv2(x.data(), v.data());
    for (size_t i = 0; i < N; ++i) r[i] = b[i] - 1.0 * v[i];
    r0 = r;
    p = r;
    std::fill(v.begin(), v.end(), 0.0);
    double rn0[2], bd[2], rho[2], rho_old[2] = {1.0, 1.0}, alpha[2] = {1.0, 1.0}, omega[2] = {1.0, 1.0};
    bool active[2];
    for (int c = 0; c < 2; ++c) {
        const double rr = dot_c(r, r, c);
        rn0[c] = std::sqrt(rr);
        rho[c] = rr;
        bd[c] = 1.0e-30 * rn0[c] * rn0[c] + 1.0e-300;
        info.iters[c] = 0;  info.converged[c] = true;  info.rel[c] = 0.0;
        active[c] = rn0[c] >= 1.0e-13;
    }
    auto stop = [&](int c, int it, bool ok) { active[c] = false; info.iters[c] = it; info.converged[c] = ok; };
    for (int iter = 0; iter < max_iter && (active[0] || active[1]); ++iter) {
        double cb[2] = {0, 0}, co[2] = {0, 0};
        for (int c = 0; c < 2; ++c) {
            if (active[c] && std::fabs(rho[c]) < bd[c]) stop(c, iter, false);
            if (active[c] && iter > 0) {
                cb[c] = (rho[c] / rho_old[c]) * (alpha[c] / (omega[c] + 1.0e-30));
                co[c] = omega[c];
            }
        }
        if (!(active[0] || active[1])) break;
        if (iter > 0)
#pragma omp parallel for schedule(static) if (n > 16384)
            for (size_t i = 0; i < N; ++i) {
                const int c = i >= (size_t)n;
                p[i] = r[i] + cb[c] * (p[i] - co[c] * v[i]);
            }
        ilu.apply(p.data(), ph.data(), 2);
        spmv2(ph.data(), v.data());
        double o0[2] = {dot_c(r0, v, 0), dot_c(r0, v, 1)};
        double ca[2] = {0, 0};
        for (int c = 0; c < 2; ++c) {
            if (active[c] && std::fabs(o0[c]) < bd[c]) stop(c, iter, false);
            if (active[c]) { alpha[c] = rho[c] / o0[c]; ca[c] = alpha[c]; }
        }
        if (!(active[0] || active[1])) break;
#pragma omp parallel for schedule(static) if (n > 16384)
        for (size_t i = 0; i < N; ++i) s[i] = r[i] - ca[i >= (size_t)n] * v[i];
        ilu.apply(s.data(), sh.data(), 2);
        spmv2(sh.data(), t.data());
        const double ss[2] = {dot_c(s, s, 0), dot_c(s, s, 1)};
        const double tt[2] = {dot_c(t, t, 0), dot_c(t, t, 1)};
        const double ts[2] = {dot_c(t, s, 0), dot_c(t, s, 1)};
        double a1[2] = {0, 0}, a2[2] = {0, 0}, cw[2] = {0, 0};
        bool s_done[2] = {false, false};
        for (int c = 0; c < 2; ++c) {
            if (!active[c]) continue;
            a1[c] = alpha[c];
            const double s_norm = std::sqrt(ss[c]);
            if (s_norm < tol * rn0[c] || s_norm < 1.0e-11) { s_done[c] = true; continue; }
            omega[c] = (tt[c] > bd[c]) ? (ts[c] / tt[c]) : 0.0;
            a2[c] = omega[c];
            cw[c] = omega[c];
        }
#pragma omp parallel for schedule(static) if (n > 16384)
        for (size_t i = 0; i < N; ++i) {
            const int c = i >= (size_t)n;
            x[i] += a1[c] * ph[i] + a2[c] * sh[i];
            r[i] = s[i] - cw[c] * t[i];
        }
        for (int c = 0; c < 2; ++c)
            if (s_done[c]) stop(c, iter + 1, true);
        if (!(active[0] || active[1])) break;
        const double rr[2] = {dot_c(r, r, 0), dot_c(r, r, 1)};
        const double r0r[2] = {dot_c(r0, r, 0), dot_c(r0, r, 1)};
        for (int c = 0; c < 2; ++c) {
            if (!active[c]) continue;
            const double r_norm = std::sqrt(rr[c]);
            info.rel[c] = r_norm / rn0[c];
            if (r_norm < tol * rn0[c] || r_norm < 1.0e-11) { stop(c, iter + 1, true); continue; }
            rho_old[c] = rho[c];
            rho[c] = r0r[c];
            if (iter + 1 == max_iter) stop(c, max_iter, false);
        }
    }
    std::copy(x.begin(), x.begin() + n, xu);
    std::copy(x.begin() + n, x.end(), xv);
    return info;
}

// ---- smoothed-aggregation AMG (src/cfa/linalg.cuh:791-862, 995-1060, 1161-1431) ---------------
Csr csr_multiply(const Csr& A, const Csr& B) {
    Csr C;
    C.rows = A.rows;
    C.cols = B.cols;
    C.rp.assign(A.rows + 1, 0);
    std::vector<int> marker(B.cols, -1), cols;
    Vec acc(B.cols, 0.0);
    std::vector<std::vector<std::pair<int, double>>> out(A.rows);
    for (int i = 0; i < A.rows; ++i) {
        cols.clear();
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a) {
            const int k = A.ci[a];
            const double av = A.v[a];
            for (int b = B.rp[k]; b < B.rp[k + 1]; ++b) {
                const int j = B.ci[b];
                if (marker[j] != i) { marker[j] = i; acc[j] = 0.0; cols.push_back(j); }
                acc[j] += av * B.v[b];
            }
        }
        std::sort(cols.begin(), cols.end());
        for (int j : cols) out[i].push_back({j, acc[j]});
        C.rp[i + 1] = C.rp[i] + (int)cols.size();
    }
    C.ci.resize(C.rp[A.rows]);
    C.v.resize(C.rp[A.rows]);
    for (int i = 0; i < A.rows; ++i) {
        int o = C.rp[i];
        for (const auto& e : out[i]) { C.ci[o] = e.first; C.v[o] = e.second; ++o; }
    }
    return C;
}

Csr csr_transpose(const Csr& A) {
    Csr T;
    T.rows = A.cols;
    T.cols = A.rows;
    T.rp.assign(A.cols + 1, 0);
    for (int c : A.ci) T.rp[c + 1]++;
    for (int i = 0; i < A.cols; ++i) T.rp[i + 1] += T.rp[i];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int> fill(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.rows; ++i)
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a) {
            const int k = fill[A.ci[a]]++;
            T.ci[k] = i;
            T.v[k] = A.v[a];
        }
    return T;
}

// 20 power iterations on D^-1 A, norms summed per 4096-row chunk and the chunks in order
double spectral_radius_DinvA(const Csr& A, const Vec& dinv) {
    const int n = A.rows, chunk = 4096, nch = (n + chunk - 1) / chunk;
    Vec x(n), y(n), py(nch), px(nch);
    for (int i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * std::sin(1.0 + i);
    double lambda = 1.0;
    for (int it = 0; it < 20; ++it) {
        for (int c = 0; c < nch; ++c) {
            double sy = 0.0, sx = 0.0;
            for (int i = c * chunk; i < std::min(n, (c + 1) * chunk); ++i) {
                double s = 0.0;
                for (int a = A.rp[i]; a < A.rp[i + 1]; ++a) s += A.v[a] * x[A.ci[a]];
                y[i] = dinv[i] * s;
                sy += y[i] * y[i];
                sx += x[i] * x[i];
            }
            py[c] = sy;
            px[c] = sx;
        }
        double nrm = 0.0, xn = 0.0;
        for (int c = 0; c < nch; ++c) { nrm += py[c]; xn += px[c]; }
        nrm = std::sqrt(nrm);
        lambda = nrm / std::sqrt(xn);
        for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
    }
    return std::max(lambda, 1.0e-12) * 1.05;
}

Csr smoothed_prolongator(const Csr& A, const Vec& dinv, double omega, double theta,
                         std::vector<int>* agg_out) {
    const int n = A.rows;
    Vec diag(n);
    for (int i = 0; i < n; ++i) diag[i] = 1.0 / dinv[i];
    std::vector<char> strong(A.ci.size());
    for (int i = 0; i < n; ++i)
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a) {
            const int j = A.ci[a];
            strong[a] = j != i && std::fabs(A.v[a]) >= theta * std::sqrt(std::fabs(diag[i] * diag[j]));
        }
    // phase 1: seed aggregates on free strong neighbourhoods (node order)
    std::vector<int> agg(n, -1);
    int na = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        bool free_nb = true;
        for (int a = A.rp[i]; a < A.rp[i + 1] && free_nb; ++a)
            if (strong[a] && agg[A.ci[a]] >= 0) free_nb = false;
        if (!free_nb) continue;
        agg[i] = na;
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a)
            if (strong[a]) agg[A.ci[a]] = na;
        ++na;
    }
    // phase 2: join the first strongly connected aggregate of phase 1
    std::vector<int> agg2(agg);
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a)
            if (strong[a] && agg[A.ci[a]] >= 0) { agg2[i] = agg[A.ci[a]]; break; }
    }
    agg.swap(agg2);
    // phase 3: remaining nodes seed new aggregates with their unassigned strong neighbours
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = na;
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a)
            if (strong[a] && agg[A.ci[a]] < 0) agg[A.ci[a]] = na;
        ++na;
    }
    if (agg_out) *agg_out = agg;
    std::vector<int> size(na, 0);
    for (int i = 0; i < n; ++i) size[agg[i]]++;
    Csr P0;
    P0.rows = n;
    P0.cols = na;
    P0.rp.resize(n + 1);
    P0.ci.resize(n);
    P0.v.resize(n);
    for (int i = 0; i < n; ++i) { P0.rp[i] = i; P0.ci[i] = agg[i]; P0.v[i] = 1.0 / std::sqrt((double)size[agg[i]]); }
    P0.rp[n] = n;
    const Csr AP0 = csr_multiply(A, P0);
    Csr P;
    P.rows = n;
    P.cols = na;
    P.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        bool own = false;
        const int oc = agg[i];
        for (int a = AP0.rp[i]; a < AP0.rp[i + 1]; ++a) {
            const int j = AP0.ci[a];
            if (!own && j > oc) { P.ci.push_back(oc); P.v.push_back(P0.v[i]); own = true; }
            double val = -omega * dinv[i] * AP0.v[a];
            if (j == oc) { val += P0.v[i]; own = true; }
            if (val != 0.0) { P.ci.push_back(j); P.v.push_back(val); }
        }
        if (!own) { P.ci.push_back(oc); P.v.push_back(P0.v[i]); }
        P.rp[i + 1] = (int)P.ci.size();
    }
    return P;
}

Vec dense_spd_inverse(Vec M, int n) {
    for (int j = 0; j < n; ++j) {
        double d = M[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= M[(size_t)j * n + k] * M[(size_t)j * n + k];
        if (!(d > 0.0)) {
            std::fprintf(stderr, "[AMG] coarse operator not SPD (pivot %.3e), aborting\n", d);
            std::exit(EXIT_FAILURE);
        }
        const double ljj = std::sqrt(d);
        M[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double s = M[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) s -= M[(size_t)i * n + k] * M[(size_t)j * n + k];
            M[(size_t)i * n + j] = s / ljj;
        }
    }
    Vec inv((size_t)n * n, 0.0), col(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) {
            double s = (i == c) ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= M[(size_t)i * n + k] * col[k];
            col[i] = s / M[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = col[i];
            for (int k = i + 1; k < n; ++k) s -= M[(size_t)k * n + i] * col[k];
            col[i] = s / M[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = col[i];
    }
    return inv;
}

void Amg::setup(const Csr& A0) {
    levels.clear();
    Csr A = A0;
    while (A.rows > coarse_target && (int)levels.size() < max_levels - 1) {
        Vec dinv(A.rows);
        for (int i = 0; i < A.rows; ++i) {
            double d = 0.0;
            for (int a = A.rp[i]; a < A.rp[i + 1]; ++a)
                if (A.ci[a] == i) d = A.v[a];
            dinv[i] = (std::fabs(d) > 1.0e-300) ? 1.0 / d : 1.0;
        }
        const double omega = 4.0 / (3.0 * spectral_radius_DinvA(A, dinv));
        Csr P = smoothed_prolongator(A, dinv, omega, theta, nullptr);
        if (P.cols >= (int)(0.85 * A.rows) || P.cols == 0) break;
        Csr R = csr_transpose(P);
        Csr Ac = csr_multiply(R, csr_multiply(A, P));
        Level L;
        L.A = std::move(A);
        L.P = std::move(P);
        L.R = std::move(R);
        L.dinv = std::move(dinv);
        L.omega = omega;
        const int nl = L.A.rows;
        L.x.assign(nl, 0.0);  L.b.assign(nl, 0.0);  L.r.assign(nl, 0.0);  L.Ax.assign(nl, 0.0);
        levels.push_back(std::move(L));
        A = std::move(Ac);
    }
    coarse_n = A.rows;
    Vec D((size_t)coarse_n * coarse_n, 0.0);
    for (int i = 0; i < coarse_n; ++i)
        for (int a = A.rp[i]; a < A.rp[i + 1]; ++a) D[(size_t)i * coarse_n + A.ci[a]] = A.v[a];
    coarse_inv = dense_spd_inverse(D, coarse_n);
    cb.assign(coarse_n, 0.0);
    cx.assign(coarse_n, 0.0);
}

namespace {
void dense_matvec(int n, const Vec& Ainv, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (n > 128)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += Ainv[(size_t)i * n + j] * x[j];
        y[i] = s;
    }
}
// y = x + omega D^-1 (b - A x)
void jacobi(const Csr& A, const Vec& dinv, double omega, const double* x, const double* b, double* y) {
#pragma omp parallel for schedule(static) if (A.rows > 16384)
    for (int i = 0; i < A.rows; ++i) {
        double s = 0.0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[i] = x[i] + omega * dinv[i] * (b[i] - s);
    }
}
}  // namespace

void Amg::vcycle(int l, const double* b) {
    Level& L = levels[l];
    const int n = L.A.rows;
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; ++i) L.x[i] = L.omega * L.dinv[i] * b[i];
    for (int s = 1; s < smooth_steps; ++s) {
        jacobi(L.A, L.dinv, L.omega, L.x.data(), b, L.Ax.data());
        std::swap(L.x, L.Ax);
    }
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = L.A.rp[i]; k < L.A.rp[i + 1]; ++k) s += L.A.v[k] * L.x[L.A.ci[k]];
        L.r[i] = b[i] - s;
    }
    const bool deeper = l + 1 < (int)levels.size();
    double* cbv = deeper ? levels[l + 1].b.data() : cb.data();
    spmv(L.R.rows, L.R.rp.data(), L.R.ci.data(), L.R.v.data(), L.r.data(), cbv);
    const double* cxv;
    if (deeper) {
        vcycle(l + 1, levels[l + 1].b.data());
        cxv = levels[l + 1].x.data();
    } else {
        dense_matvec(coarse_n, coarse_inv, cb.data(), cx.data());
        cxv = cx.data();
    }
#pragma omp parallel for schedule(static) if (n > 16384)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = L.P.rp[i]; k < L.P.rp[i + 1]; ++k) s += L.P.v[k] * cxv[L.P.ci[k]];
        L.x[i] += s;
    }
    for (int s = 0; s < smooth_steps; ++s) {
        jacobi(L.A, L.dinv, L.omega, L.x.data(), b, L.Ax.data());
        std::swap(L.x, L.Ax);
    }
}

void Amg::apply(const double* r, double* z) {
    if (levels.empty()) {
        dense_matvec(coarse_n, coarse_inv, r, z);
        return;
    }
    vcycle(0, r);
    std::copy(levels[0].x.begin(), levels[0].x.end(), z);
}

// ---- AMG remap (src/cfa/linalg.cuh:1074-1157): the level-0 prolongator carried to a new graph --
// rp/ci: the new operator's graph; new_to_old[i]: the old row of a kept node, -1 for a new node.
// New nodes get, pass after pass, the mean of the rows of their neighbours that had one before the
// pass (entries gathered in neighbour order, sorted by column with std::sort as the reference
// does, equal columns summed in that order). *ok = false when a pass sets nothing.
Csr amg_remap_p0(const Csr& P0, int n, const int* rp, const int* ci, const std::vector<int>& new_to_old, bool* ok) {
    Csr P;
    P.rows = n;
    P.cols = P0.cols;
    P.rp.assign(n + 1, 0);
    std::vector<int> extra_of(n, -1);
    std::vector<std::vector<std::pair<int, double>>> extra;
    std::vector<int> pending;
    for (int i = 0; i < n; ++i)
        if (new_to_old[i] < 0) pending.push_back(i);
    std::vector<char> has_row(n);
    for (int i = 0; i < n; ++i) has_row[i] = new_to_old[i] >= 0;
    *ok = true;
    while (!pending.empty()) {
        std::vector<int> still, set_now;
        for (int i : pending) {
            std::vector<std::pair<int, double>> acc;
            int cnt = 0;
            for (int a = rp[i]; a < rp[i + 1]; ++a) {
                const int k = ci[a];
                if (k == i || !has_row[k]) continue;
                ++cnt;
                if (new_to_old[k] >= 0) {
                    for (int b = P0.rp[new_to_old[k]]; b < P0.rp[new_to_old[k] + 1]; ++b)
                        acc.push_back({P0.ci[b], P0.v[b]});
                } else {
                    acc.insert(acc.end(), extra[extra_of[k]].begin(), extra[extra_of[k]].end());
                }
            }
            if (cnt == 0) {
                still.push_back(i);
                continue;
            }
            std::sort(acc.begin(), acc.end(),
                      [](const std::pair<int, double>& x, const std::pair<int, double>& y) { return x.first < y.first; });
            std::vector<std::pair<int, double>> row;
            for (const auto& e : acc) {
                if (!row.empty() && row.back().first == e.first) row.back().second += e.second;
                else row.push_back(e);
            }
            for (auto& e : row) e.second /= cnt;
            extra_of[i] = (int)extra.size();
            extra.push_back(row);
            set_now.push_back(i);
        }
        if (set_now.empty()) {
            *ok = false;
            return P;
        }
        for (int i : set_now) has_row[i] = 1;
        pending.swap(still);
    }
    for (int i = 0; i < n; ++i) {
        const int j = new_to_old[i];
        P.rp[i + 1] = P.rp[i] + (j >= 0 ? P0.rp[j + 1] - P0.rp[j] : (int)extra[extra_of[i]].size());
    }
    for (int i = 0; i < n; ++i) {
        const int j = new_to_old[i];
        if (j >= 0) {
            P.ci.insert(P.ci.end(), P0.ci.begin() + P0.rp[j], P0.ci.begin() + P0.rp[j + 1]);
            P.v.insert(P.v.end(), P0.v.begin() + P0.rp[j], P0.v.begin() + P0.rp[j + 1]);
        } else {
            for (const auto& e : extra[extra_of[i]]) {
                P.ci.push_back(e.first);
                P.v.push_back(e.second);
            }
        }
    }
    return P;
}

}  // namespace oracle
