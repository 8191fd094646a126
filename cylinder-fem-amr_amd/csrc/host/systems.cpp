// Solver-study dumps (--dump-systems N): the reader of the files the program writes at step N
// (<prefix>_momentum.bin, src/cylinder_fem_amr.cu:331-352 and :375-378; <prefix>_pressure.bin,
// :426-438) and solves of the dumped systems with the step's own solvers, as the reference's
// study programs use them (scripts/momentum_study/mb.cu: SpMV and multicolour-ILU timings and
// the ILU-BiCGSTAB solve of the momentum system; mb_amg.cu: SA-AMG set-up, per-level timings and
// PCG on K_p alone). The files hold no OSS operator, so the pressure solve is on K_p only.
#include <cmath>
#include <cstdio>
#include <string>

#include "device.hpp"

namespace cfa {
namespace {

struct DumpedSystem {
    int n = 0, nnz = 0;
    std::vector<int> rp, ci, tag;
    std::vector<double> val, b_u, b_v, x_u, x_v, cx, cy, sol_u, sol_v;   // b_v.. only for momentum
    bool has_solution = false;
};

// Reads `count` items of T; false when the file ends first.
template <typename T>
bool take(std::FILE* f, std::vector<T>& out, size_t count) {
    out.resize(count);
    return count == 0 || std::fread(out.data(), sizeof(T), count, f) == count;
}

// The pattern must be a CSR of an n x n matrix: row_ptr from 0 to nnz, non-decreasing, columns
// in range. Everything downstream (colouring, ILU split, AMG set-up) indexes with it.
bool valid_pattern(const DumpedSystem& s) {
    if (s.n <= 0 || s.nnz < 0 || s.rp.front() != 0 || s.rp.back() != s.nnz) return false;
    for (int i = 0; i < s.n; ++i)
        if (s.rp[i + 1] < s.rp[i]) return false;
    for (int c : s.ci)
        if (c < 0 || c >= s.n) return false;
    return true;
}

long file_size(std::FILE* f) {
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    return sz;
}

// momentum file: n, nnz; rp[n+1], ci[nnz]; val[nnz], rhs_u, rhs_v, u0, v0, x, y [n]; bc_tag[n];
// optionally u, v [n] (written after the solve). The size must be one of the two layouts.
int read_dump(const char* path, bool momentum, DumpedSystem& s) {
    if (!path) return CFA_E_ARG;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return CFA_E_IO;
    const long size = file_size(f);
    int hdr[2] = {0, 0};
    int rc = std::fread(hdr, sizeof(int), 2, f) == 2 ? CFA_OK : CFA_E_IO;
    s.n = hdr[0];
    s.nnz = hdr[1];
    if (rc == CFA_OK && (s.n <= 0 || s.nnz <= 0)) rc = CFA_E_IO;
    if (rc == CFA_OK) {
        const long n = s.n, nnz = s.nnz;
        const long base = 8 + 4 * (n + 1) + 12 * nnz;
        const long mom = base + 6 * 8 * n + 4 * n;
        if (momentum) {
            s.has_solution = size == mom + 16 * n;
            if (size != mom && !s.has_solution) rc = CFA_E_IO;
        } else if (size != base + 16 * n) {
            rc = CFA_E_IO;
        }
    }
    if (rc == CFA_OK) {
        const size_t n = (size_t)s.n, nnz = (size_t)s.nnz;
        bool ok = take(f, s.rp, n + 1) && take(f, s.ci, nnz) && take(f, s.val, nnz) && take(f, s.b_u, n);
        if (momentum) {
            ok = ok && take(f, s.b_v, n) && take(f, s.x_u, n) && take(f, s.x_v, n) && take(f, s.cx, n) &&
                 take(f, s.cy, n) && take(f, s.tag, n);
            if (s.has_solution) ok = ok && take(f, s.sol_u, n) && take(f, s.sol_v, n);
        } else {
            ok = ok && take(f, s.x_u, n);
        }
        if (!ok || !valid_pattern(s)) rc = CFA_E_IO;
    }
    std::fclose(f);
    return rc;
}

double rel_max_diff(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0.0, m = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        d = std::max(d, std::fabs(a[i] - b[i]));
        m = std::max(m, std::fabs(b[i]));
    }
    return m > 0.0 ? d / m : d;
}

// Algorithmic bytes of one V(2,2) cycle: per level three smoother-class passes over A, the
// residual, restriction and prolongation, plus the dense coarse solve (FP64 values and vectors).
double vcycle_bytes(const AmgPreconditioner& amg) {
    double b = 8.0 * amg.coarse_n * (double)amg.coarse_n;
    for (const auto& L : amg.levels) {
        const double n = L->A.rows, a = L->A.nnz;
        b += 3.0 * (12.0 * a + 44.0 * n) + (12.0 * a + 36.0 * n) + (12.0 * L->R.nnz + 28.0 * n) +
             (12.0 * L->P.nnz + 36.0 * n);
    }
    return b;
}

double synced_ms(Clock::time_point t0) {
    dev_sync();
    return ms_since(t0);
}

}  // namespace
}  // namespace cfa

using namespace cfa;

extern "C" {

int cfa_dump_header(const char* path, int* n, int* nnz) {
    if (!path || !n || !nnz) return CFA_E_ARG;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return CFA_E_IO;
    int hdr[2] = {0, 0};
    const bool ok = std::fread(hdr, sizeof(int), 2, f) == 2;
    std::fclose(f);
    if (!ok || hdr[0] <= 0 || hdr[1] <= 0) return CFA_E_IO;
    *n = hdr[0];
    *nnz = hdr[1];
    return CFA_OK;
}

// BiCGSTAB2 with multicolour ILU(0) on a dumped momentum system, from the dumped initial guess.
int cfa_dumped_momentum_solve(const char* path, double tol, int max_iter, int precision, int kernel_reps,
                              double* x_u, double* x_v, double* out) {
    if (!out || !(tol > 0.0) || max_iter < 1 || (precision != 0 && precision != 1) || kernel_reps < 0)
        return CFA_E_ARG;
    DumpedSystem s;
    if (const int rc = read_dump(path, true, s)) return rc;
    GPU_CHECK(cfa_init(0));
    kernel_timings().clear();
    const int n = s.n, nnz = s.nnz;
    DevArray<int> rp, ci;
    DevArray<double> val, bu, bv, xu, xv;
    rp.upload(s.rp);
    ci.upload(s.ci);
    val.upload(s.val);
    bu.upload(s.b_u);
    bv.upload(s.b_v);
    xu.upload(s.x_u);
    xv.upload(s.x_v);
    MulticolourIlu0 ilu;
    ilu.keep_fp32 = precision == 1;
    auto t0 = Clock::now();
    ilu.set_pattern(n, s.rp, s.ci, rp, ci);
    const double colour_ms = synced_ms(t0);
    t0 = Clock::now();
    ilu.factor(val);
    const double factor_ms = synced_ms(t0);
    GpuBicgstab2 sv;
    sv.precond_fp32 = precision == 1;
    t0 = Clock::now();
    sv.solve(n, rp, ci, val, bu, bv, xu, xv, max_iter, tol, ilu);
    const double solve_ms = synced_ms(t0);
    const std::vector<double> ru = xu.to_host(n), rv = xv.to_host(n);
    if (x_u) std::copy(ru.begin(), ru.end(), x_u);
    if (x_v) std::copy(rv.begin(), rv.end(), x_v);
    double dev = -1.0;
    if (s.has_solution) dev = std::max(rel_max_diff(ru, s.sol_u), rel_max_diff(rv, s.sol_v));
    if (kernel_reps > 0) {
        DevArray<double> t_in, t_out;
        t_in.reserve(2 * (size_t)n);
        t_out.reserve(2 * (size_t)n);
        dev_copy(t_in, bu, n);
        dev_copy(t_in + n, bv, n);
        const double dn = n, dnnz = nnz;
        kernel_timings().push_back({"cfa_spmv, 2 right-hand sides", 12.0 * dnnz + 36.0 * dn + 4.0, time_launches(kernel_reps, [&] {
            GPU_CHECK(cfa_spmv(n, rp, ci, val, t_in, t_out, 2, solver_stream()));
        }), kernel_reps});
        kernel_timings().push_back({"ilu0 apply, 2 components (all colours)", 12.0 * dnnz + 100.0 * dn,
                                    time_launches(kernel_reps, [&] { ilu.apply(t_in, t_out, 2); }), kernel_reps});
    }
    const double v[11] = {(double)n, (double)nnz, (double)ilu.n_colours, colour_ms, factor_ms, solve_ms,
                          (double)sv.iters[0], (double)sv.iters[1], sv.converged[0] ? 1.0 : 0.0,
                          sv.converged[1] ? 1.0 : 0.0, dev};
    for (int k = 0; k < 11; ++k) out[k] = v[k];
    return CFA_OK;
}

// SA-AMG PCG on K_p of a dumped pressure system, from the dumped initial guess (from_zero = 0) or
// from x = 0 (as scripts/momentum_study/mb_amg.cu does).
int cfa_dumped_pressure_solve(const char* path, double tol, int max_iter, int precision, int from_zero,
                              int kernel_reps, double* x, double* out) {
    if (!out || !(tol > 0.0) || max_iter < 1 || (precision != 0 && precision != 1) || kernel_reps < 0)
        return CFA_E_ARG;
    DumpedSystem s;
    if (const int rc = read_dump(path, false, s)) return rc;
    GPU_CHECK(cfa_init(0));
    kernel_timings().clear();
    const int n = s.n, nnz = s.nnz;
    DevArray<int> rp, ci;
    DevArray<double> val, b, xd;
    rp.upload(s.rp);
    ci.upload(s.ci);
    val.upload(s.val);
    b.upload(s.b_u);
    xd.upload(s.x_u);
    if (from_zero) dev_zero(xd, n * sizeof(double));
    AmgPreconditioner amg;
    amg.fp32 = precision == 1;
    auto t0 = Clock::now();
    amg.setup(n, nnz, rp, ci, val);
    const double setup_ms = synced_ms(t0);
    GpuLinearSolver solver;
    t0 = Clock::now();
    const int iters = solver.solve_pcg_fused(n, rp, ci, val, b, xd, max_iter, tol, nullptr, &amg);
    const double solve_ms = synced_ms(t0);
    if (x) xd.download(x, n);
    if (kernel_reps > 0) {
        DevArray<double> t_out;
        t_out.reserve(n);
        const double dn = n, dnnz = nnz;
        kernel_timings().push_back({"cfa_spmv, 1 right-hand side", 12.0 * dnnz + 20.0 * dn + 4.0, time_launches(kernel_reps, [&] {
            GPU_CHECK(cfa_spmv(n, rp, ci, val, xd, t_out, 1, solver_stream()));
        }), kernel_reps});
        kernel_timings().push_back({"V-cycle", vcycle_bytes(amg), time_launches(kernel_reps, [&] { amg.apply(b, t_out); }),
                                    kernel_reps});
    }
    const double v[8] = {(double)n, (double)nnz, (double)amg.num_levels(), setup_ms, solve_ms, (double)iters,
                         solver.converged ? 1.0 : 0.0, solver.last_rel_residual};
    for (int k = 0; k < 8; ++k) out[k] = v[k];
    return CFA_OK;
}

}  // extern "C"
