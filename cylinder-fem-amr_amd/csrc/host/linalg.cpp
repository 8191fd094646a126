// Solver classes over the C ABI. The smoothed-aggregation set-up (src/cfa/linalg.cuh:995-1060 and
// 1276-1431: strength |a_ij| >= theta sqrt(|a_ii a_jj|) with theta 0.08, three-phase greedy
// aggregation in row order, tentative prolongator 1/sqrt(|aggregate|), one damped Jacobi
// smoothing step with omega = 4 / (3 * 1.05 * rho(D^-1 A)) from 20 power iterations, Galerkin
// R A P with R = P^T, stop at <= 300 rows, 15 levels or when P.cols >= 0.85 rows, dense Cholesky
// inverse of the coarsest operator) is sequenced here level by level over the device kernels of
// amg_setup.hip and aggregate.hip; the host inverts the coarse operator and, for remap(), fills
// the prolongator rows of new nodes layer by layer.
#include "linalg.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cfa {

// ---------------------------------------------------------------------------------------------
KrylovWorkspace::~KrylovWorkspace() {
    cfa_krylov_ws_release(&ws_);
    if (status_host_) cfa_host_free(status_host_);
}

cfa_krylov_ws_t* KrylovWorkspace::get(int n) {
    if (!status_host_) {
        void* d_alias = nullptr;
        GPU_CHECK(cfa_host_malloc(&status_host_, &d_alias, 16 * sizeof(int)));
        std::memset(status_host_, 0, 16 * sizeof(int));
        ws_.status_host = static_cast<int*>(status_host_);
        ws_.status_dev = static_cast<int*>(d_alias);
        part_.reserve(8 * 4096, 0);
        scal_.reserve(64, 0);
        ws_.part = part_.ptr;
        ws_.scal = scal_.ptr;
        const char* env = std::getenv("CFA_KRYLOV_GRAPHS");
        ws_.use_graphs = (env && env[0] == '0') ? 0 : 1;
    }
    if (n > ws_.cap) {
        const int cap = n + n / 2 + 256;
        for (int k = 0; k < 12; ++k) {
            vec_[k].release();
            vec_[k].reserve(2 * (size_t)cap, 0);
            ws_.vec[k] = vec_[k].ptr;
        }
        ws_.cap = cap;
    }
    return &ws_;
}

// ---------------------------------------------------------------------------------------------
// GpuLinearSolver (src/cfa/linalg.cuh:120-348)
// ---------------------------------------------------------------------------------------------
namespace {
void call_op(const double* in, double* out, void* ctx) {
    (*static_cast<const std::function<void(const double*, double*)>*>(ctx))(in, out);
}
}  // namespace

void GpuLinearSolver::allocate(int n) {
    ws_.get(n);
    max_n = ws_.capacity();
}

double GpuLinearSolver::dot(int n, const double* d_x, const double* d_y) {
    dot_ws_.reserve(2048, 0);
    double r = 0.0;
    GPU_CHECK(cfa_dot(n, d_x, d_y, &r, dot_ws_.ptr, solver_stream()));
    return r;
}

int GpuLinearSolver::solve_pcg(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                               const double* d_b, double* d_x, int max_iter, double tol,
                               const std::function<void(const double*, double*)>* extra_op,
                               const std::function<void(const double*, double*)>* precond) {
    cfa_solve_result_t res{};
    GPU_CHECK(cfa_pcg(n, d_row_ptr, d_col_ind, d_val, d_b, d_x, max_iter, tol, nullptr,
                      extra_op ? call_op : nullptr, const_cast<void*>(static_cast<const void*>(extra_op)), nullptr,
                      precond ? call_op : nullptr, const_cast<void*>(static_cast<const void*>(precond)), ws_.get(n),
                      &res, solver_stream()));
    max_n = ws_.capacity();
    converged = res.converged[0] != 0;
    if (res.rel_residual[0] >= 0.0) last_rel_residual = res.rel_residual[0];
    last_iters = res.iters[0];
    return res.iters[0];
}

int GpuLinearSolver::solve_pcg_fused(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                                     const double* d_b, double* d_x, int max_iter, double tol,
                                     const cfa_oss_op_t* oss, AmgPreconditioner* amg) {
    cfa_solve_result_t res{};
    GPU_CHECK(cfa_pcg(n, d_row_ptr, d_col_ind, d_val, d_b, d_x, max_iter, tol, oss, nullptr, nullptr,
                      amg ? amg->view() : nullptr, nullptr, nullptr, ws_.get(n), &res, solver_stream()));
    max_n = ws_.capacity();
    converged = res.converged[0] != 0;
    if (res.rel_residual[0] >= 0.0) last_rel_residual = res.rel_residual[0];
    last_iters = res.iters[0];
    return res.iters[0];
}

int GpuLinearSolver::solve_pcg_fused(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                                     const double* d_b, double* d_x, int max_iter, double tol,
                                     const cfa_oss_op_t* oss, const cfa_amg_t* amg_view, cfa_stream_t stream) {
    cfa_solve_result_t res{};
    GPU_CHECK(cfa_pcg(n, d_row_ptr, d_col_ind, d_val, d_b, d_x, max_iter, tol, oss, nullptr, nullptr, amg_view,
                      nullptr, nullptr, ws_.get(n), &res, stream));
    max_n = ws_.capacity();
    converged = res.converged[0] != 0;
    if (res.rel_residual[0] >= 0.0) last_rel_residual = res.rel_residual[0];
    last_iters = res.iters[0];
    return res.iters[0];
}

int GpuLinearSolver::solve_bicgstab(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                                    const double* d_b, double* d_x, int max_iter, double tol) {
    cfa_solve_result_t res{};
    GPU_CHECK(cfa_bicgstab_jacobi(n, d_row_ptr, d_col_ind, d_val, d_b, d_x, max_iter, tol, ws_.get(n), &res,
                                  solver_stream()));
    max_n = ws_.capacity();
    converged = res.converged[0] != 0;
    last_rel_residual = res.rel_residual[0];
    last_iters = res.iters[0];
    return res.iters[0];
}

// ---------------------------------------------------------------------------------------------
// MulticolourIlu0 (src/cfa/linalg.cuh:435-568)
// ---------------------------------------------------------------------------------------------
// Elimination schedule of ILU(0) in colour order, row by row in colour-row order (the loop order
// of src/cfa/linalg.cuh:375-391): for c < colour(i), the L entries a_ik of colour c in CSR order;
// for each, the entries of row k whose column j is in row i's pattern with colour(j) > c, in row
// k's order.
void ilu_elimination_schedule(int n, const int* rp, const int* ci, const std::vector<int>& colour,
                              const std::vector<int>& rows, const std::vector<int>& diag, std::vector<int>& el_ptr,
                              std::vector<int>& el, std::vector<int>& upd) {
    std::vector<int> up_cnt(n + 1, 0);
    el_ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i)
        if (diag[i] < 0) fatal("[ILU] row without a diagonal entry");
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t) {
        const int i = rows[t];
        int steps = 0, ups = 0;
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int k = ci[a];
            if (k == i || colour[k] >= colour[i]) continue;
            ++steps;
            for (int m = rp[k]; m < rp[k + 1]; ++m) {
                const int j = ci[m];
                if (colour[j] <= colour[k]) continue;
                for (int b = rp[i]; b < rp[i + 1]; ++b)
                    if (ci[b] == j) { ++ups; break; }
            }
        }
        el_ptr[t + 1] = steps;
        up_cnt[t + 1] = ups;
    }
    for (int t = 0; t < n; ++t) { el_ptr[t + 1] += el_ptr[t]; up_cnt[t + 1] += up_cnt[t]; }
    el.assign(4 * (size_t)std::max(el_ptr[n], 1), 0);
    upd.assign(2 * (size_t)std::max(up_cnt[n], 1), 0);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t) {
        const int i = rows[t];
        int e = el_ptr[t], u = up_cnt[t];
        for (int c = 0; c < colour[i]; ++c)
            for (int a = rp[i]; a < rp[i + 1]; ++a) {
                const int k = ci[a];
                if (k == i || colour[k] != c) continue;
                el[4 * e] = a;
                el[4 * e + 1] = diag[k];
                el[4 * e + 2] = u;
                for (int m = rp[k]; m < rp[k + 1]; ++m) {
                    const int j = ci[m];
                    if (colour[j] <= c) continue;
                    for (int b = rp[i]; b < rp[i + 1]; ++b)
                        if (ci[b] == j) { upd[2 * u] = b; upd[2 * u + 1] = m; ++u; break; }
                }
                el[4 * e + 3] = u;
                ++e;
            }
    }
}

namespace {
// Sliced-ELL layout of split rows (row pointer ptr over colour-order rows, columns col) colour by
// colour: colour c's rows take positions pos0[c] + t (pos0[c] a multiple of 256); inside each
// 256-position window the rows are ordered by decreasing length, ties by row; 64-position slices
// hold their slots column-major, as many per position as the slice's first (longest) row has.
// row: the colour's row at each position (-1 past its last); src: the split entry of each slot
// (-1 for padding, whose column is 0).
struct SellRows {
    std::vector<int> row, sp, col, src;
    int longest = 0;
};
SellRows sell_rows(const std::vector<int>& cs, const std::vector<int>& pos0, const std::vector<int>& ptr,
                   const std::vector<int>& col) {
    SellRows S;
    const int nc = (int)cs.size() - 1, total = pos0[nc];
    S.row.assign(total, -1);
    std::vector<int> order;
    for (int c = 0; c < nc; ++c) {
        const int m = cs[c + 1] - cs[c];
        auto len = [&](int t) { return ptr[cs[c] + t + 1] - ptr[cs[c] + t]; };
        for (int w0 = 0; w0 < m; w0 += 256) {
            order.clear();
            for (int t = w0; t < std::min(m, w0 + 256); ++t) order.push_back(t);
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len(a) > len(b); });
            for (size_t k = 0; k < order.size(); ++k) S.row[pos0[c] + w0 + k] = order[k];
        }
    }
    const int slices = total / 64;
    S.sp.assign(slices + 1, 0);
    for (int q = 0; q < slices; ++q) {
        int L = 0;
        for (int l = 0; l < 64; ++l) {
            const int p = 64 * q + l, t = S.row[p];
            if (t < 0) continue;
            const int c = (int)(std::upper_bound(pos0.begin(), pos0.end(), p) - pos0.begin()) - 1;
            L = std::max(L, ptr[cs[c] + t + 1] - ptr[cs[c] + t]);
        }
        S.sp[q + 1] = S.sp[q] + 64 * L;
        S.longest = std::max(S.longest, L);
    }
    S.col.assign(S.sp[slices], 0);
    S.src.assign(S.sp[slices], -1);
    for (int c = 0; c < nc; ++c)
        for (int p = pos0[c]; p < pos0[c + 1]; ++p) {
            const int t = S.row[p];
            if (t < 0) continue;
            const int i = cs[c] + t, base = S.sp[p / 64] + p % 64;
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
                S.col[base + 64 * (k - ptr[i])] = col[k];
                S.src[base + 64 * (k - ptr[i])] = k;
            }
        }
    return S;
}
}  // namespace

MulticolourIlu0::Pattern MulticolourIlu0::build_pattern(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                                                        int sell_min_rows) {
    const auto t0 = Clock::now();
    Pattern P;
    P.n = n;
    P.nnz = rp[n];
    // greedy colouring in node order: the smallest colour not used by an earlier neighbour
    P.colour.assign(n, -1);
    P.diag.assign(n, -1);
    std::vector<int> mark;
    std::vector<int>& colour = P.colour;
    for (int i = 0; i < n; ++i) {
        if (rp[i + 1] - rp[i] > 32) {
            std::fprintf(stderr, "[ILU] row %d has %d entries (limit %d)\n", i, rp[i + 1] - rp[i], 32);
            std::exit(EXIT_FAILURE);
        }
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int c = colour[ci[a]];
            if (c >= 0) mark[c] = i;
            if (ci[a] == i) P.diag[i] = a;
        }
        int c = 0;
        while (c < P.n_colours && mark[c] == i) ++c;
        if (c == P.n_colours) { ++P.n_colours; mark.push_back(-1); }
        colour[i] = c;
    }
    P.colour_start.assign(P.n_colours + 1, 0);
    for (int i = 0; i < n; ++i) P.colour_start[colour[i] + 1]++;
    for (int c = 0; c < P.n_colours; ++c) P.colour_start[c + 1] += P.colour_start[c];
    P.rows.resize(n);
    P.pos.resize(n);
    std::vector<int> fill(P.colour_start.begin(), P.colour_start.end() - 1);
    for (int i = 0; i < n; ++i) {
        P.pos[i] = fill[colour[i]]++;
        P.rows[P.pos[i]] = i;
    }
    // split pattern in colour-row order: L = earlier colours, U = later colours
    P.lp.assign(n + 1, 0);
    P.up.assign(n + 1, 0);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t) {
        const int i = P.rows[t];
        int l = 0, u = 0;
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            if (ci[a] == i) continue;
            if (colour[ci[a]] < colour[i]) ++l; else ++u;
        }
        P.lp[t + 1] = l;
        P.up[t + 1] = u;
    }
    for (int t = 0; t < n; ++t) { P.lp[t + 1] += P.lp[t]; P.up[t + 1] += P.up[t]; }
    P.nl = P.lp[n];
    P.nu = P.up[n];
    P.lc.resize(P.nl);
    P.uc.resize(P.nu);
    P.l_src.resize(P.nl);
    P.u_src.resize(P.nu);
    P.d_src.resize(n);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; ++t) {
        const int i = P.rows[t];
        int l = P.lp[t], u = P.up[t];
        for (int a = rp[i]; a < rp[i + 1]; ++a) {
            const int j = ci[a];
            if (j == i) P.d_src[t] = a;
            else if (colour[j] < colour[i]) { P.lc[l] = P.pos[j]; P.l_src[l++] = a; }
            else { P.uc[u] = P.pos[j]; P.u_src[u++] = a; }
        }
    }
    ilu_elimination_schedule(n, rp.data(), ci.data(), colour, P.rows, P.diag, P.el_ptr, P.el, P.upd);
    const char* sell_env = std::getenv("CFA_SELL");
    if (n > 0 && n >= sell_min_rows && P.n_colours > 0 && !(sell_env && sell_env[0] == '0')) {
        P.s_pos0.assign(P.n_colours + 1, 0);
        for (int c = 0; c < P.n_colours; ++c)
            P.s_pos0[c + 1] = P.s_pos0[c] + (P.colour_start[c + 1] - P.colour_start[c] + 255) / 256 * 256;
        SellRows L = sell_rows(P.colour_start, P.s_pos0, P.lp, P.lc), U = sell_rows(P.colour_start, P.s_pos0, P.up, P.uc);
        P.s_lrow = std::move(L.row); P.s_lsp = std::move(L.sp); P.s_lc = std::move(L.col); P.s_lsrc = std::move(L.src);
        P.s_urow = std::move(U.row); P.s_usp = std::move(U.sp); P.s_uc = std::move(U.col); P.s_usrc = std::move(U.src);
        P.s_llen = L.longest;
        P.s_ulen = U.longest;
    }
    P.build_ms = ms_since(t0);
    return P;
}

void MulticolourIlu0::set_pattern(int n_, const std::vector<int>& rp, const std::vector<int>& ci, const int* d_rp,
                                  const int* d_ci) {
    set_pattern(build_pattern(n_, rp, ci, sell_min_rows), d_rp, d_ci);
}


void MulticolourIlu0::set_pattern(const Pattern& P, const int* d_rp, const int* d_ci) {
    const auto t0 = Clock::now();
    n = P.n;
    nnz = P.nnz;
    n_colours = P.n_colours;
    nl = P.nl;
    nu = P.nu;
    colour_start = P.colour_start;
    d_row_ptr = d_rp;
    d_col_ind = d_ci;
    d_el_ptr.upload(P.el_ptr);
    d_el.upload(P.el);
    d_upd.upload(P.upd);
    d_rows.upload(P.rows);
    d_colour.upload(P.colour);
    d_diag_pos.upload(P.diag);
    d_lp.upload(P.lp);
    d_up.upload(P.up);
    d_lc.upload(P.lc);
    d_uc.upload(P.uc);
    d_l_src.upload(P.l_src);
    d_u_src.upload(P.u_src);
    d_d_src.upload(P.d_src);
    d_pos.upload(P.pos);
    d_r_perm.reserve(2 * (size_t)n, 64);
    d_z_perm.reserve(2 * (size_t)n, 64);
    d_lu.reserve(nnz, 256);
    d_lv.reserve(std::max(nl, 1), 256);
    d_uv.reserve(std::max(nu, 1), 256);
    d_dinv.reserve(n, 64);
    if (keep_fp32) {
        d_lv32.reserve(std::max(nl, 1), 256);
        d_uv32.reserve(std::max(nu, 1), 256);
        d_dinv32.reserve(n, 64);
    }
    // the sliced-ELL copies of the split rows for the one-lane sweeps
    has_sell = false;
    const char* env = std::getenv("CFA_SELL");
    if (!P.s_pos0.empty() && !(env && env[0] == '0')) {
        if (P.s_llen <= CFA_SELL_MAX_LEN && P.s_ulen <= CFA_SELL_MAX_LEN) {
            s_pos0 = P.s_pos0;
            s_llen = P.s_llen;
            s_ulen = P.s_ulen;
            s_lslots = (int)P.s_lc.size();
            s_uslots = (int)P.s_uc.size();
            // a copy without slots still gets one-element arrays, so that its pointers are set
            auto put = [](DevArray<int>& d, const std::vector<int>& h, int fill) {
                if (h.empty()) d.upload(std::vector<int>(1, fill));
                else d.upload(h);
            };
            s_lrow.upload(P.s_lrow);
            s_lsp.upload(P.s_lsp);
            put(s_lc, P.s_lc, 0);
            put(s_lsrc, P.s_lsrc, -1);
            s_urow.upload(P.s_urow);
            s_usp.upload(P.s_usp);
            put(s_uc, P.s_uc, 0);
            put(s_usrc, P.s_usrc, -1);
            s_lv.reserve(std::max(s_lslots, 1), 256);
            s_uv.reserve(std::max(s_uslots, 1), 256);
            if (keep_fp32) {
                s_lv32.reserve(std::max(s_lslots, 1), 256);
                s_uv32.reserve(std::max(s_uslots, 1), 256);
            }
            has_sell = true;
        }
    }
    refresh_view();
    colour_ms += P.build_ms + ms_since(t0);
}

void MulticolourIlu0::refresh_view() {
    view_.n = n;
    view_.nnz = nnz;
    view_.n_colours = n_colours;
    view_.nl = nl;
    view_.nu = nu;
    view_.colour_start = colour_start.data();
    view_.rows = d_rows;
    view_.colour = d_colour;
    view_.diag_pos = d_diag_pos;
    view_.row_ptr = d_row_ptr;
    view_.col_ind = d_col_ind;
    view_.lu = d_lu;
    view_.lp = d_lp;
    view_.up = d_up;
    view_.lc = d_lc;
    view_.uc = d_uc;
    view_.pos = d_pos;
    view_.r_perm = d_r_perm;
    view_.z_perm = d_z_perm;
    view_.l_src = d_l_src;
    view_.u_src = d_u_src;
    view_.d_src = d_d_src;
    view_.lv = d_lv;
    view_.uv = d_uv;
    view_.dinv = d_dinv;
    view_.lv32 = keep_fp32 ? d_lv32.ptr : nullptr;
    view_.uv32 = keep_fp32 ? d_uv32.ptr : nullptr;
    view_.dinv32 = keep_fp32 ? d_dinv32.ptr : nullptr;
    view_.el_ptr = d_el_ptr;
    view_.el = d_el;
    view_.upd = d_upd;
    view_.s_pos0 = has_sell ? s_pos0.data() : nullptr;
    view_.s_lrow = has_sell ? s_lrow.ptr : nullptr;
    view_.s_lsp = has_sell ? s_lsp.ptr : nullptr;
    view_.s_lc = has_sell ? s_lc.ptr : nullptr;
    view_.s_lsrc = has_sell ? s_lsrc.ptr : nullptr;
    view_.s_lv = has_sell ? s_lv.ptr : nullptr;
    view_.s_lv32 = has_sell && keep_fp32 ? s_lv32.ptr : nullptr;
    view_.s_urow = has_sell ? s_urow.ptr : nullptr;
    view_.s_usp = has_sell ? s_usp.ptr : nullptr;
    view_.s_uc = has_sell ? s_uc.ptr : nullptr;
    view_.s_usrc = has_sell ? s_usrc.ptr : nullptr;
    view_.s_uv = has_sell ? s_uv.ptr : nullptr;
    view_.s_uv32 = has_sell && keep_fp32 ? s_uv32.ptr : nullptr;
    view_.s_llen = has_sell ? s_llen : 0;
    view_.s_ulen = has_sell ? s_ulen : 0;
    view_.s_lslots = has_sell ? s_lslots : 0;
    view_.s_uslots = has_sell ? s_uslots : 0;
}

void MulticolourIlu0::factor(const double* d_val) { GPU_CHECK(cfa_ilu0_factor(&view_, d_val, solver_stream())); }

void MulticolourIlu0::apply(const double* r, double* z, int nrhs) const {
    GPU_CHECK(cfa_ilu0_apply(&view_, r, z, nrhs, solver_stream()));
}

// ---------------------------------------------------------------------------------------------
// GpuBicgstab2 (src/cfa/linalg.cuh:642-779)
// ---------------------------------------------------------------------------------------------
void GpuBicgstab2::solve(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val, const double* b_u,
                         const double* b_v, double* x_u, double* x_v, int max_iter, double tol,
                         const MulticolourIlu0& ilu) {
    cfa_solve_result_t res{};
    GPU_CHECK(cfa_bicgstab2(n, d_row_ptr, d_col_ind, d_val, b_u, b_v, x_u, x_v, max_iter, tol, ilu.view(),
                            precond_fp32 ? 1 : 0, ws_.get(n), &res, solver_stream()));
    for (int c = 0; c < 2; ++c) {
        iters[c] = res.iters[c];
        converged[c] = res.converged[c] != 0;
        last_rel_residual[c] = res.rel_residual[c];
    }
}


// ---------------------------------------------------------------------------------------------
// SA-AMG (src/cfa/linalg.cuh:939-1432): the set-up runs on the device level by level
// ---------------------------------------------------------------------------------------------
int csr_lanes_per_row(int rows, long long nnz) {
    // Measured on MI355X (one Jacobi sweep per level, lanes 1..32, uniform level-4 and level-5
    // pressure hierarchies): rows of ~7 entries run fastest with 2-4 lanes (one lane is 2.2x
    // slower on 460k rows), ~15 entries with 4-8, and short levels (< 20k rows) are latency bound
    // and run fastest when a group covers the whole row. So: about four entries per lane, at
    // least 2 lanes; one lane per entry (up to 32) below 20k rows.
    const double per_row = rows > 0 ? (double)nnz / rows : 1.0;
    const double target = rows < 20000 ? per_row : per_row / 4.0;
    int tpr = 2;
    while (tpr < 32 && tpr < target) tpr *= 2;
    return tpr;
}

void AmgPreconditioner::upload_csr_from_device(int rows, int nnz, const int* rp, const int* ci, const double* v,
                                               DevCsr& D) {
    D.rows = rows;
    D.nnz = nnz;
    D.tpr = csr_lanes_per_row(rows, nnz);
    D.rp.reserve(rows + 1);
    D.ci.reserve(std::max(nnz, 1));
    D.v.reserve(std::max(nnz, 1));
    GPU_CHECK(cfa_memcpy(D.rp.ptr, rp, (rows + 1) * sizeof(int), 2, solver_stream()));
    if (nnz) {
        GPU_CHECK(cfa_memcpy(D.ci.ptr, ci, nnz * sizeof(int), 2, solver_stream()));
        GPU_CHECK(cfa_memcpy(D.v.ptr, v, nnz * sizeof(double), 2, solver_stream()));
    }
}

void AmgPreconditioner::ensure_level_vectors(Level& L, int n) {
    L.dinv.reserve(n, 16);
    L.x.reserve(n, 16);
    L.b.reserve(n, 16);
    L.r.reserve(n, 16);
    L.Ax.reserve(n, 16);
}

void AmgPreconditioner::AuxStream::ensure() {
    if (s) return;
    GPU_CHECK(cfa_stream_create(&s));
    GPU_CHECK(cfa_timer_create(&ready));
    GPU_CHECK(cfa_timer_create(&done));
}

AmgPreconditioner::AuxStream::~AuxStream() {
    if (!s) return;
    (void)cfa_stream_sync(s);
    (void)cfa_timer_destroy(ready);
    (void)cfa_timer_destroy(done);
    (void)cfa_stream_destroy(s);
}

void AmgPreconditioner::release() {
    for (auto& L : levels) spare_.push_back(std::move(L));
    levels.clear();
    coarse_n = 0;
}

// C = A B on the device (host csr_multiply results, bit for bit); false when a row has more
// products than the device kernel holds
bool AmgPreconditioner::device_product(int rows, int cols, int a_nnz, const int* arp, const int* aci,
                                       const double* av, const int* brp, const int* bci, const double* bv, DevCsr& C,
                                       int left_max_row) {
    const cfa_stream_t s = solver_stream();
    d_cnt_.reserve(std::max(rows, 1));
    d_flag_.reserve(2);
    const size_t sws = cfa_scan_counts_ws_bytes(rows);
    d_sws_.reserve(std::max<size_t>(sws, 1));
    C.rp.reserve(rows + 1);
    // method 0: one thread per row merging the right rows (rows of A up to 16 entries);
    // method 2: one wavefront per row merging them in LDS (up to 64 entries, 1024 products);
    // method 1: one workgroup per row sorting the products in LDS (up to 4096 per row).
    // The count pass, the scan and one download of {overflow flag, total} per attempt; a method
    // whose limit a row exceeds hands the product to the next.
    std::vector<int> methods;
    if (!force_sorted_products) {
        if (left_max_row >= 0 && left_max_row <= 16) methods.push_back(0);
        if (left_max_row < 0 || left_max_row <= 64) methods.push_back(2);
        methods.push_back(1);
    }
    for (const int method : methods) {
        dev_zero(d_flag_, 2 * sizeof(int));
        GPU_CHECK(cfa_spgemm(rows, arp, aci, av, brp, bci, bv, 0, method, d_cnt_, nullptr, nullptr, nullptr, d_flag_,
                             s));
        GPU_CHECK(cfa_scan_counts(rows, d_cnt_, C.rp, d_flag_.ptr + 1, d_sws_, sws, s));
        int ft[2] = {0, 0};
        d_flag_.download(ft, 2);
        if (ft[0]) continue;
        const int total = ft[1];
        C.ci.reserve(std::max(total, 1));
        C.v.reserve(std::max(total, 1));
        GPU_CHECK(cfa_spgemm(rows, arp, aci, av, brp, bci, bv, 1, method, nullptr, C.rp, C.ci, C.v, d_flag_, s));
        C.rows = rows;
        C.nnz = total;
        C.tpr = csr_lanes_per_row(rows, total);
        return true;
    }
    // every product listed in global memory, stably sorted by (row, column), summed per key
    long long products = 0, entries = 0;
    size_t ws = cfa_spgemm_sorted_ws_bytes(rows, cols, a_nnz, 0);
    if (ws == 0) return false;
    d_gws_.reserve(ws);
    GPU_CHECK(cfa_spgemm_sorted(rows, cols, a_nnz, arp, aci, av, brp, bci, bv, 0, 0, d_gws_, ws, &products, nullptr,
                                nullptr, nullptr, s));
    ws = cfa_spgemm_sorted_ws_bytes(rows, cols, a_nnz, products);
    if (ws == 0) return false;
    d_gws_.reserve(ws);
    GPU_CHECK(cfa_spgemm_sorted(rows, cols, a_nnz, arp, aci, av, brp, bci, bv, 1, products, d_gws_, ws, &entries,
                                nullptr, nullptr, nullptr, s));
    C.ci.reserve(std::max<long long>(entries, 1));
    C.v.reserve(std::max<long long>(entries, 1));
    GPU_CHECK(cfa_spgemm_sorted(rows, cols, a_nnz, arp, aci, av, brp, bci, bv, 2, products, d_gws_, ws, nullptr, C.rp,
                                C.ci, C.v, s));
    C.rows = rows;
    C.nnz = (int)entries;
    C.tpr = csr_lanes_per_row(rows, C.nnz);
    ++sorted_products;
    return true;
}


namespace {
std::vector<int> down_ints(const int* d, size_t n) {
    std::vector<int> h(n);
    if (n) GPU_CHECK(cfa_memcpy(h.data(), d, n * sizeof(int), 1, solver_stream()));
    return h;
}

// Inverse of a small SPD matrix (the coarse level, at most a few hundred rows): Cholesky factor
// L (row by row, in place in the lower triangle), W = L^-1 by forward substitution per column,
// then M^-1 = W^T W. The reference inverts the same matrix through its Cholesky factor
// (src/cfa/linalg.cuh:1395-1430); an indefinite pivot ends the run as there.
std::vector<double> spd_inverse(std::vector<double> M, int n) {
    auto at = [&](int i, int j) -> double& { return M[(size_t)i * n + j]; };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = at(i, j);
            for (int k = 0; k < j; ++k) s -= at(i, k) * at(j, k);
            if (i == j) {
                if (!(s > 0.0)) {
                    std::fprintf(stderr, "[AMG] coarse operator not SPD (pivot %.3e), aborting\n", s);
                    std::exit(EXIT_FAILURE);
                }
                at(i, i) = std::sqrt(s);
            } else {
                at(i, j) = s / at(j, j);
            }
        }
    std::vector<double> W((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
    for (int c = 0; c < n; ++c) {
        W[(size_t)c * n + c] = 1.0 / at(c, c);
        for (int i = c + 1; i < n; ++i) {
            double s = 0.0;
            for (int k = c; k < i; ++k) s += at(i, k) * W[(size_t)k * n + c];
            W[(size_t)i * n + c] = -s / at(i, i);
        }
    }
    std::vector<double> inv((size_t)n * n);
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
            for (int k = i; k < n; ++k) s += W[(size_t)k * n + i] * W[(size_t)k * n + j];
            inv[(size_t)i * n + j] = inv[(size_t)j * n + i] = s;
        }
    return inv;
}
}  // namespace

// One level on the device: D^-1, the spectral radius of D^-1 A (power iteration), the
// three-phase aggregation, P = P0 - omega D^-1 A P0, R = P^T and the Galerkin product R (A P),
// which becomes the next level's operator (dev_next_). Returns 1 with the next level's size, 0
// when coarsening stalls (aggregates >= 0.85 rows or none).
int AmgPreconditioner::device_level(int n, int nnz, const int* arp, const int* aci, const double* av,
                                    std::vector<std::unique_ptr<Level>>& pool, int& n_next, int& nnz_next) {
    const cfa_stream_t s = solver_stream();
    const bool first = levels.empty();
    phase_start();
    std::unique_ptr<Level> L;
    if (!pool.empty()) {
        L = std::move(pool.front());
        pool.erase(pool.begin());
    } else {
        L = std::make_unique<Level>();
    }
    auto give_back = [&]() { pool.insert(pool.begin(), std::move(L)); };
    ensure_level_vectors(*L, n);
    GPU_CHECK(cfa_amg_dinv(n, arp, aci, av, L->dinv, s));
    if (x0_dev_n_ < n) {   // power-iteration start vector 1 + 0.1 sin(1 + i) (linalg.cuh:1279)
        // its entries do not depend on n: extend it, with headroom for the meshes that follow
        const int old_n = (int)x0_host_.size();
        const int new_n = std::max(n, old_n + old_n / 2);
        x0_host_.resize(new_n);
        for (int i = old_n; i < new_n; ++i) x0_host_[i] = 1.0 + 0.1 * std::sin(1.0 + i);
        d_x0_.upload(x0_host_);
        x0_dev_n_ = new_n;
    }
    const size_t pw = cfa_amg_power_ws_doubles(n);
    d_pw_.reserve(pw);
    // the power iteration on the second stream, beside the aggregation below
    aux_.ensure();
    GPU_CHECK(cfa_timer_record(aux_.ready, s));
    GPU_CHECK(cfa_stream_wait(aux_.s, aux_.ready));
    GPU_CHECK(cfa_amg_power(n, arp, aci, av, L->dinv, d_x0_, 20, d_pw_, nullptr, aux_.s));
    GPU_CHECK(cfa_timer_record(aux_.done, aux_.s));
    phase_end(PH_RHO);
    // aggregation; then one download: {aggregates, phase-1 aggregates, longest row, rounds} and lambda
    const size_t aws = cfa_amg_aggregate_ws_bytes(n, nnz);
    d_aws_.reserve(std::max<size_t>(aws, 1));
    d_agg_.reserve(n);
    d_p0v_.reserve(n);
    d_info_.reserve(4);
    GPU_CHECK(cfa_amg_aggregate(n, nnz, arp, aci, av, L->dinv, theta, d_aws_, aws, d_agg_, d_p0v_, d_info_, s));
    int info[4] = {0, 0, 0, 0};
    d_info_.download(info, 4);
    double lambda = 0.0;
    GPU_CHECK(cfa_stream_wait(s, aux_.done));
    GPU_CHECK(cfa_memcpy(&lambda, d_pw_.ptr + pw - 1, sizeof(double), 1, s));
    phase_end(PH_AGG);
    const int na = info[0], left_max = info[2];
    front_rounds += info[3];
    const double omega = 4.0 / (3.0 * (std::max(lambda, 1.0e-12) * 1.05));
    if (na >= (int)(0.85 * n) || na == 0) {
        give_back();
        return 0;
    }
    if (first && keep_aggregates) last_aggregates = d_agg_.to_host(n);
    if (iota_dev_n_ < n + 1) {   // row pointer of P0 (one entry per row), with headroom
        const int m = std::max(n + 1, iota_dev_n_ + iota_dev_n_ / 2);
        std::vector<int> iota(m);
        for (int i = 0; i < m; ++i) iota[i] = i;
        d_iota_.upload(iota);
        iota_dev_n_ = m;
    }
    // P = P0 - omega D^-1 A P0
    if (!device_product(n, na, nnz, arp, aci, av, d_iota_, d_agg_, d_p0v_, dev_tmp_, left_max)) {
        give_back();
        return -1;
    }
    d_cnt_.reserve(std::max(std::max(n, na), 1));
    GPU_CHECK(cfa_amg_prolongator(n, dev_tmp_.rp, dev_tmp_.ci, dev_tmp_.v, d_agg_, d_p0v_, L->dinv, omega, 0, d_cnt_,
                                  nullptr, nullptr, nullptr, s));
    L->P.rp.reserve(n + 1);
    const size_t sws = cfa_scan_counts_ws_bytes(n);
    d_sws_.reserve(std::max<size_t>(sws, 1));
    d_flag_.reserve(2);
    GPU_CHECK(cfa_scan_counts(n, d_cnt_, L->P.rp, d_flag_.ptr + 1, d_sws_, sws, s));
    int p_nnz = 0;
    GPU_CHECK(cfa_memcpy(&p_nnz, d_flag_.ptr + 1, sizeof(int), 1, s));
    L->P.ci.reserve(std::max(p_nnz, 1));
    L->P.v.reserve(std::max(p_nnz, 1));
    GPU_CHECK(cfa_amg_prolongator(n, dev_tmp_.rp, dev_tmp_.ci, dev_tmp_.v, d_agg_, d_p0v_, L->dinv, omega, 1, nullptr,
                                  L->P.rp, L->P.ci, L->P.v, s));
    L->P.rows = n;
    L->P.nnz = p_nnz;
    L->P.tpr = csr_lanes_per_row(n, p_nnz);
    phase_end(PH_SMOOTH);
    // R = P^T
    L->R.rp.reserve(na + 1);
    L->R.ci.reserve(std::max(p_nnz, 1));
    L->R.v.reserve(std::max(p_nnz, 1));
    const size_t tws = cfa_csr_transpose_ws_bytes(n, na, p_nnz);
    d_tws_.reserve(std::max<size_t>(tws, 1));
    GPU_CHECK(cfa_csr_transpose(n, na, p_nnz, L->P.rp, L->P.ci, L->P.v, d_tws_, tws, L->R.rp, L->R.ci, L->R.v, s));
    L->R.rows = na;
    L->R.nnz = p_nnz;
    L->R.tpr = csr_lanes_per_row(na, p_nnz);
    phase_end(PH_TRANSPOSE);
    // Galerkin product R (A P) (rows of R cover an aggregate's neighbourhood: the LDS-sort kernel)
    if (!device_product(n, na, nnz, arp, aci, av, L->P.rp, L->P.ci, L->P.v, dev_tmp_, left_max) ||
        !device_product(na, na, p_nnz, L->R.rp, L->R.ci, L->R.v, dev_tmp_.rp, dev_tmp_.ci, dev_tmp_.v, dev_next2_,
                        -1)) {
        give_back();
        return -1;
    }
    phase_end(PH_GALERKIN);
    // the level's operator: the caller's matrix on level 0, the previous product after that
    if (first) upload_csr_from_device(n, nnz, arp, aci, av, L->A);
    else std::swap(L->A, dev_next_);
    L->A.tpr = csr_lanes_per_row(n, L->A.nnz);
    L->omega = omega;
    if (first && keep_p0_host) {   // remap() reads the level-0 prolongator on the host
        P0_host.rows = n;
        P0_host.cols = na;
        P0_host.rp = down_ints(L->P.rp, n + 1);
        P0_host.ci = down_ints(L->P.ci, p_nnz);
        P0_host.v = L->P.v.to_host(p_nnz);
    }
    std::swap(dev_next_, dev_next2_);
    levels.push_back(std::move(L));
    n_next = na;
    nnz_next = dev_next_.nnz;
    phase_end(PH_DOWNLOAD);
    return 1;
}

void AmgPreconditioner::setup(int n, int nnz, const int* d_row_ptr, const int* d_col_ind, const double* d_val) {
    const auto t0 = Clock::now();
    // keep the buffers of the previous hierarchy for reuse, finest first
    std::vector<std::unique_ptr<Level>> old;
    old.swap(levels);
    for (auto it = spare_.begin(); it != spare_.end(); ++it) old.push_back(std::move(*it));
    spare_.clear();
    last_aggregates.clear();
    device_levels = 0;
    front_rounds = 0;
    sorted_products = 0;
    // the current level's operator: the caller's matrix, then the Galerkin products
    const int* arp = d_row_ptr;
    const int* aci = d_col_ind;
    const double* av = d_val;
    int rows = n, entries = nnz;
    while (rows > coarse_target && (int)levels.size() < max_levels - 1) {
        int n_next = 0, nnz_next = 0;
        const int r = device_level(rows, entries, arp, aci, av, old, n_next, nnz_next);
        if (r == 0) break;
        if (r < 0) fatal("[AMG] a sparse product of the set-up exceeds 2^31 - 1 products; the hierarchy cannot be built");
        ++device_levels;
        arp = dev_next_.rp;
        aci = dev_next_.ci;
        av = dev_next_.v;
        rows = n_next;
        entries = nnz_next;
    }
    for (auto& L : old) spare_.push_back(std::move(L));

    // coarse operator to the host, dense inverse back to the device
    phase_start();
    coarse_n = rows;
    const std::vector<int> crp = down_ints(arp, rows + 1), cci = down_ints(aci, entries);
    std::vector<double> cv(entries);
    if (entries) GPU_CHECK(cfa_memcpy(cv.data(), av, entries * sizeof(double), 1, solver_stream()));
    std::vector<double> D((size_t)coarse_n * coarse_n, 0.0);
    for (int i = 0; i < coarse_n; ++i)
        for (int a = crp[i]; a < crp[i + 1]; ++a) D[(size_t)i * coarse_n + cci[a]] = cv[a];
    const std::vector<double> inv = spd_inverse(D, coarse_n);
    d_coarse_inv.upload(inv);
    d_cx.reserve(std::max(coarse_n, 1), 16);
    d_cb.reserve(std::max(coarse_n, 1), 16);
    refresh_view();
    dev_sync();
    phase_end(PH_COARSE);
    setup_ms = ms_since(t0);
}


// The sliced-ELL copy of a level's operator (CFA_SELL=0 keeps the CSR rows): every entry load of a
// sweep's wavefront is one contiguous run, and the sums keep the CSR kernels' order for the level's
// lanes per row (1, 2 or 4). Levels with a row longer than CFA_SELL_MAX_LEN keep the CSR kernels.
void AmgPreconditioner::build_sell(Level& L) {
    L.has_sell = false;
    const int n = L.A.rows;
    const char* env = std::getenv("CFA_SELL");
    if (n <= 0 || (env && env[0] == '0') || (L.A.tpr != 1 && L.A.tpr != 2 && L.A.tpr != 4)) return;
    const cfa_stream_t s = solver_stream();
    const int slices = (n + 63) / 64;
    L.s_row.reserve((size_t)slices * 64);
    L.s_sp.reserve((size_t)slices + 1);
    const size_t wsb = cfa_sell_ws_bytes(n);
    if (wsb == 0) return;
    d_sell_ws_.reserve(wsb, 0);
    d_sell_info_.reserve(2, 0);
    GPU_CHECK(cfa_sell_plan(n, L.A.rp, L.s_row, L.s_sp, d_sell_info_, d_sell_ws_, wsb, s));
    int info[2] = {0, 0};
    d_sell_info_.download(info, 2);
    if (info[1] > CFA_SELL_MAX_LEN) return;
    L.s_entries = info[0];
    L.s_len = info[1];
    L.s_ci.reserve(std::max(L.s_entries, 1), 256);
    L.s_v.reserve(std::max(L.s_entries, 1), 256);
    GPU_CHECK(cfa_sell_fill(n, L.A.rp, L.A.ci, L.A.v, L.s_row, L.s_sp, L.s_ci, L.s_v, s));
    L.has_sell = true;
}

void AmgPreconditioner::refresh_view() {
    for (size_t l = 0; l < levels.size(); ++l) {   // level 0 only: the coarser levels' rows are longer
        if (l == 0) build_sell(*levels[l]);
        else levels[l]->has_sell = false;
    }
    desc_.resize(levels.size());
    for (size_t l = 0; l < levels.size(); ++l) {
        const Level& L = *levels[l];
        cfa_amg_level_t& d = desc_[l];
        d.rows = L.A.rows;
        d.cols = L.P.rows > 0 ? (int)(L.R.rows) : 0;
        d.a_nnz = L.A.nnz;
        d.p_nnz = L.P.nnz;
        d.r_nnz = L.R.nnz;
        d.tpr_a = L.A.tpr;
        d.tpr_p = L.P.tpr;
        d.tpr_r = L.R.tpr;
        d.a_rp = L.A.rp; d.a_ci = L.A.ci; d.a_v = L.A.v;
        d.p_rp = L.P.rp; d.p_ci = L.P.ci; d.p_v = L.P.v;
        d.r_rp = L.R.rp; d.r_ci = L.R.ci; d.r_v = L.R.v;
        d.dinv = L.dinv;
        d.omega = L.omega;
        d.x = L.x; d.b = L.b; d.r = L.r; d.ax = L.Ax;
        d.a_v32 = d.p_v32 = d.r_v32 = d.dinv32 = nullptr;
        d.x32 = d.b32 = d.r32 = d.ax32 = nullptr;
        d.s_len = L.has_sell ? L.s_len : 0;
        d.s_sp = L.has_sell ? L.s_sp.ptr : nullptr;
        d.s_row = L.has_sell ? L.s_row.ptr : nullptr;
        d.s_ci = L.has_sell ? L.s_ci.ptr : nullptr;
        d.s_v = L.has_sell ? L.s_v.ptr : nullptr;
        d.s_v32 = nullptr;
    }
    if (fp32) {   // FP32 copies of the values and FP32 work vectors
        const cfa_stream_t s = solver_stream();
        for (size_t l = 0; l < levels.size(); ++l) {
            Level& L = *levels[l];
            cfa_amg_level_t& d = desc_[l];
            const int n = L.A.rows;
            L.A32.reserve(std::max(L.A.nnz, 1), 256);
            L.P32.reserve(std::max(L.P.nnz, 1), 256);
            L.R32.reserve(std::max(L.R.nnz, 1), 256);
            for (DevArray<float>* v : {&L.dinv32, &L.x32, &L.b32, &L.r32, &L.Ax32}) v->reserve(n, 16);
            GPU_CHECK(cfa_to_f32(L.A.nnz, L.A.v, L.A32, s));
            GPU_CHECK(cfa_to_f32(L.P.nnz, L.P.v, L.P32, s));
            GPU_CHECK(cfa_to_f32(L.R.nnz, L.R.v, L.R32, s));
            GPU_CHECK(cfa_to_f32(n, L.dinv, L.dinv32, s));
            d.a_v32 = L.A32; d.p_v32 = L.P32; d.r_v32 = L.R32; d.dinv32 = L.dinv32;
            d.x32 = L.x32; d.b32 = L.b32; d.r32 = L.r32; d.ax32 = L.Ax32;
            if (L.has_sell) {
                L.s_v32.reserve(std::max(L.s_entries, 1), 256);
                GPU_CHECK(cfa_to_f32(L.s_entries, L.s_v, L.s_v32, s));
                d.s_v32 = L.s_v32;
            }
        }
        d_cx32.reserve(std::max(coarse_n, 1), 16);
        d_cb32.reserve(std::max(coarse_n, 1), 16);
    }
    view_.use_f32 = fp32 ? 1 : 0;
    view_.cb32 = fp32 ? d_cb32.ptr : nullptr;
    view_.cx32 = fp32 ? d_cx32.ptr : nullptr;
    view_.n_levels = (int)levels.size();
    view_.levels = desc_.data();
    view_.coarse_n = coarse_n;
    view_.coarse_inv = d_coarse_inv;
    view_.cb = d_cb;
    view_.cx = d_cx;
    view_.smooth_steps = smooth_steps;
}

const cfa_amg_t* AmgPreconditioner::second_view() {
    desc2_ = desc_;
    view2_ = view_;
    while (vec2_.size() < levels.size()) vec2_.push_back(std::make_unique<WorkVectors>());
    for (size_t l = 0; l < levels.size(); ++l) {
        WorkVectors& w = *vec2_[l];
        cfa_amg_level_t& d = desc2_[l];
        const int n = d.rows;
        for (DevArray<double>* v : {&w.x, &w.b, &w.r, &w.Ax}) v->reserve(std::max(n, 1), 16);
        d.x = w.x; d.b = w.b; d.r = w.r; d.ax = w.Ax;
        if (fp32) {
            for (DevArray<float>* v : {&w.x32, &w.b32, &w.r32, &w.Ax32}) v->reserve(std::max(n, 1), 16);
            d.x32 = w.x32; d.b32 = w.b32; d.r32 = w.r32; d.ax32 = w.Ax32;
        }
    }
    d_cx2_.reserve(std::max(coarse_n, 1), 16);
    d_cb2_.reserve(std::max(coarse_n, 1), 16);
    view2_.cx = d_cx2_;
    view2_.cb = d_cb2_;
    if (fp32) {
        d_cx32_2_.reserve(std::max(coarse_n, 1), 16);
        d_cb32_2_.reserve(std::max(coarse_n, 1), 16);
        view2_.cx32 = d_cx32_2_;
        view2_.cb32 = d_cb32_2_;
    }
    view2_.levels = desc2_.data();
    return &view2_;
}

void AmgPreconditioner::apply(const double* d_r, double* d_z) {
    GPU_CHECK(cfa_amg_apply(&view_, d_r, d_z, solver_stream()));
}

// Carries the level-0 prolongator to a new mesh (src/cfa/linalg.cuh:1074-1157): kept nodes keep
// their row, a new node gets the mean of the rows of its already-set graph neighbours, pass by
// pass; coarser levels, omega and the coarse inverse stay.

// Carries the level-0 prolongator to a new mesh instead of a new set-up (remap, the reference's
// src/cfa/linalg.cuh:1074-1157, off by default): a node kept by the remesh (new_to_old[i] >= 0)
// keeps its old row; the new nodes are filled layer by layer outwards from the kept ones, each
// taking the mean of the rows of its graph neighbours filled in earlier layers (columns merged,
// equal columns summed in neighbour order). Coarser levels, omega and the coarse inverse stay;
// level 0 gets the new operator, its inverse diagonal, P and R = P^T (transposed on the device).
// A new node that no layer reaches (no kept node in its part of the graph) means a new set-up.
void AmgPreconditioner::remap(int n, int nnz, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                              const int* h_rp, const int* h_ci, const std::vector<int>& new_to_old) {
    if (levels.empty() || P0_host.rows == 0) {
        setup(n, nnz, d_row_ptr, d_col_ind, d_val);
        return;
    }
    const auto t0 = Clock::now();
    phase_start();
    const HostCsr& Po = P0_host;
    // layer of every node: 0 kept, k filled from layers < k, -1 not reached yet
    std::vector<int> layer(n, -1), frontier, next;
    for (int i = 0; i < n; ++i)
        if (new_to_old[i] >= 0) layer[i] = 0;
    for (int i = 0; i < n; ++i)
        if (layer[i] < 0) frontier.push_back(i);
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    auto row_of = [&](int k, std::vector<std::pair<int, double>>& out) {
        if (new_to_old[k] >= 0) {
            const int j = new_to_old[k];
            for (int b = Po.rp[j]; b < Po.rp[j + 1]; ++b) out.push_back({Po.ci[b], Po.v[b]});
        } else {
            out.insert(out.end(), rows[k].begin(), rows[k].end());
        }
    };
    for (int depth = 1; !frontier.empty(); ++depth) {
        next.clear();
        std::vector<int> filled;
        for (int i : frontier) {
            std::vector<std::pair<int, double>> acc;
            int used = 0;
            for (int a = h_rp[i]; a < h_rp[i + 1]; ++a) {
                const int k = h_ci[a];
                if (k == i || layer[k] < 0 || layer[k] >= depth) continue;
                ++used;
                row_of(k, acc);
            }
            if (!used) { next.push_back(i); continue; }
            std::stable_sort(acc.begin(), acc.end(),
                             [](const std::pair<int, double>& x, const std::pair<int, double>& y) { return x.first < y.first; });
            auto& out = rows[i];
            for (const auto& e : acc) {
                if (!out.empty() && out.back().first == e.first) out.back().second += e.second;
                else out.push_back(e);
            }
            for (auto& e : out) e.second /= used;
            filled.push_back(i);
        }
        if (filled.empty()) {   // unreachable new nodes
            setup(n, nnz, d_row_ptr, d_col_ind, d_val);
            return;
        }
        for (int i : filled) layer[i] = depth;
        frontier.swap(next);
    }
    HostCsr P;
    P.rows = n;
    P.cols = Po.cols;
    P.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const int j = new_to_old[i];
        P.rp[i + 1] = P.rp[i] + (j >= 0 ? Po.rp[j + 1] - Po.rp[j] : (int)rows[i].size());
    }
    P.ci.reserve(P.rp[n]);
    P.v.reserve(P.rp[n]);
    for (int i = 0; i < n; ++i) {
        std::vector<std::pair<int, double>> r;
        row_of(i, r);
        for (const auto& e : r) { P.ci.push_back(e.first); P.v.push_back(e.second); }
    }
    const cfa_stream_t s = solver_stream();
    Level& L = *levels[0];
    upload_csr_from_device(n, nnz, d_row_ptr, d_col_ind, d_val, L.A);
    L.P.rows = n;
    L.P.nnz = P.rp[n];
    L.P.tpr = csr_lanes_per_row(n, L.P.nnz);
    L.P.rp.upload(P.rp);
    L.P.ci.upload(P.ci);
    L.P.v.upload(P.v);
    const int cols = P.cols, pn = P.rp[n];
    L.R.rows = cols;
    L.R.nnz = pn;
    L.R.tpr = csr_lanes_per_row(cols, pn);
    L.R.rp.reserve(cols + 1);
    L.R.ci.reserve(std::max(pn, 1));
    L.R.v.reserve(std::max(pn, 1));
    const size_t tws = cfa_csr_transpose_ws_bytes(n, cols, pn);
    d_tws_.reserve(std::max<size_t>(tws, 1));
    GPU_CHECK(cfa_csr_transpose(n, cols, pn, L.P.rp, L.P.ci, L.P.v, d_tws_, tws, L.R.rp, L.R.ci, L.R.v, s));
    ensure_level_vectors(L, n);
    GPU_CHECK(cfa_inv_diag(n, L.A.rp, L.A.ci, L.A.v, 1.0e-300, L.dinv, s));
    P0_host = std::move(P);
    ++remaps;
    refresh_view();
    dev_sync();
    phase_end(PH_REMAP);
    setup_ms = ms_since(t0);
}

}  // namespace cfa
