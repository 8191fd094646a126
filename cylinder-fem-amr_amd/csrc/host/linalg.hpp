// Solver classes with the reference's interfaces (src/cfa/linalg.cuh): GpuLinearSolver
// (:120-348), MulticolourIlu0 (:435-568), GpuBicgstab2 (:642-779), AmgPreconditioner
// (:939-1432). The numerical work runs in the gfx950 kernels behind include/cfa_hip.h, the
// smoothed-aggregation set-up included (the reference runs that on the host with OpenMP).
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "gpu.hpp"

namespace cfa {

// Krylov workspace shared by the solvers (vectors, partial sums, device scalars, status word)
class KrylovWorkspace {
public:
    KrylovWorkspace() = default;
    KrylovWorkspace(const KrylovWorkspace&) = delete;
    KrylovWorkspace& operator=(const KrylovWorkspace&) = delete;
    ~KrylovWorkspace();
    cfa_krylov_ws_t* get(int n);   // capacity >= n
    int capacity() const { return ws_.cap; }
private:
    cfa_krylov_ws_t ws_{};
    DevArray<double> vec_[12];
    DevArray<double> part_, scal_;
    void* status_host_ = nullptr;
};

class AmgPreconditioner;

class GpuLinearSolver {
public:
    int max_n = 0;
    bool converged = true;
    double last_rel_residual = 0.0;
    int last_iters = 0;

    void allocate(int n);
    double dot(int n, const double* d_x, const double* d_y);

    // Preconditioned CG on CSR val (+ extra_op) with precond (AMG apply etc.), else Jacobi.
    int solve_pcg(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                  const double* d_b, double* d_x, int max_iter = 400, double tol = 1.0e-7,
                  const std::function<void(const double*, double*)>* extra_op = nullptr,
                  const std::function<void(const double*, double*)>* precond = nullptr);
    // The same solve with the OSS operator and the AMG V-cycle fused into the iteration.
    int solve_pcg_fused(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                        const double* d_b, double* d_x, int max_iter, double tol,
                        const cfa_oss_op_t* oss, AmgPreconditioner* amg);
    // The same with a given hierarchy view and stream (a second solve running beside the first)
    int solve_pcg_fused(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                        const double* d_b, double* d_x, int max_iter, double tol,
                        const cfa_oss_op_t* oss, const cfa_amg_t* amg_view, cfa_stream_t stream);
    int solve_bicgstab(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
                       const double* d_b, double* d_x, int max_iter = 200, double tol = 1.0e-7);
private:
    KrylovWorkspace ws_;
    DevArray<double> dot_ws_;
};

class MulticolourIlu0 {
public:
    int n = 0, nnz = 0, n_colours = 0, nl = 0, nu = 0;
    double colour_ms = 0.0;
    std::vector<int> colour_start;
    DevArray<int> d_rows, d_colour, d_diag_pos;
    DevArray<double> d_lu;
    const int* d_row_ptr = nullptr;
    const int* d_col_ind = nullptr;
    DevArray<int> d_lp, d_up, d_lc, d_uc, d_l_src, d_u_src, d_d_src, d_pos;
    DevArray<double> d_r_perm, d_z_perm;   // the apply's vectors in colour order
    DevArray<int> d_el_ptr, d_el, d_upd;   // elimination schedule of the factorization
    DevArray<double> d_lv, d_uv, d_dinv;
    DevArray<float> d_lv32, d_uv32, d_dinv32;   // FP32 factors (mixed precision)
    bool keep_fp32 = false;

    MulticolourIlu0() = default;
    MulticolourIlu0(const MulticolourIlu0&) = delete;
    MulticolourIlu0& operator=(const MulticolourIlu0&) = delete;

    void set_pattern(int n, const std::vector<int>& rp, const std::vector<int>& ci, const int* d_rp,
                     const int* d_ci);
    // the host part of set_pattern (colouring, split pattern, elimination schedule); it touches no
    // device memory, so a remesh worker thread can prepare it
    struct Pattern {
        int n = 0, nnz = 0, n_colours = 0, nl = 0, nu = 0;
        // lc, uc: colour-order positions of the split entries' columns; pos: of each row
        std::vector<int> colour_start, rows, colour, diag, lp, up, lc, uc, l_src, u_src, d_src, el_ptr, el, upd, pos;
        // sliced-ELL layout of the split rows (cfa_ilu_t.s_*), built with the pattern
        std::vector<int> s_pos0, s_lrow, s_lsp, s_lc, s_lsrc, s_urow, s_usp, s_uc, s_usrc;
        int s_llen = 0, s_ulen = 0;
        double build_ms = 0.0;
    };
    // Below about a million rows the two-component apply's bytes (about 184 per row) stay in the
    // 256 MB Infinity Cache and the CSR sweeps are as fast: the level-6 adaptive benchmark (300k
    // rows) ran 62.9 against 63.6-64.1 steps/s with the copies
    static constexpr int SELL_MIN_ROWS = 1 << 20;
    static Pattern build_pattern(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                                 int sell_min_rows = SELL_MIN_ROWS);
    void set_pattern(const Pattern& P, const int* d_rp, const int* d_ci);
    // sliced-ELL copies of the split rows (cfa_ilu_t.s_*): built with the pattern when it has at
    // least sell_min_rows rows, every split row has at most CFA_SELL_MAX_LEN entries and CFA_SELL
    // is not 0
    int sell_min_rows = SELL_MIN_ROWS;
    bool has_sell = false;
    std::vector<int> s_pos0;
    DevArray<int> s_lrow, s_lsp, s_lc, s_lsrc, s_urow, s_usp, s_uc, s_usrc;
    DevArray<double> s_lv, s_uv;
    DevArray<float> s_lv32, s_uv32;
    int s_llen = 0, s_ulen = 0, s_lslots = 0, s_uslots = 0;
    void factor(const double* d_val);
    void apply(const double* r, double* z, int nrhs = 1) const;
    const cfa_ilu_t* view() const { return &view_; }
private:
    cfa_ilu_t view_{};
    void refresh_view();
};

// elimination schedule of the colour-ordered ILU(0) factorization (layout of cfa_ilu_t)
void ilu_elimination_schedule(int n, const int* rp, const int* ci, const std::vector<int>& colour,
                              const std::vector<int>& rows, const std::vector<int>& diag, std::vector<int>& el_ptr,
                              std::vector<int>& el, std::vector<int>& upd);

class GpuBicgstab2 {
public:
    int iters[2] = {0, 0};
    bool converged[2] = {true, true};
    double last_rel_residual[2] = {0.0, 0.0};
    bool precond_fp32 = false;   // apply the FP32 copy of the ILU factors

    GpuBicgstab2() = default;
    GpuBicgstab2(const GpuBicgstab2&) = delete;
    GpuBicgstab2& operator=(const GpuBicgstab2&) = delete;

    void solve(int n, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
               const double* b_u, const double* b_v, double* x_u, double* x_v, int max_iter,
               double tol, const MulticolourIlu0& ilu);
private:
    KrylovWorkspace ws_;
};

// Host CSR matrix (the level-0 prolongator kept for remap)
struct HostCsr {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
};

class AmgPreconditioner {
public:
    struct DevCsr {
        int rows = 0, nnz = 0, tpr = 1;
        DevArray<int> rp, ci;
        DevArray<double> v;
    };
    struct Level {
        DevCsr A, P, R;
        DevArray<double> dinv;
        double omega = 0.0;
        DevArray<double> x, b, r, Ax;
        // FP32 copies and work vectors of the mixed-precision cycle
        DevArray<float> A32, P32, R32, dinv32, x32, b32, r32, Ax32;
        // sliced-ELL copy of A that the level's sweeps and residual read (level 0; cfa_sell_plan)
        bool has_sell = false;
        int s_len = 0, s_entries = 0;
        DevArray<int> s_sp, s_row, s_ci;
        DevArray<double> s_v;
        DevArray<float> s_v32;
    };

    int smooth_steps = 2;
    double theta = 0.08;
    int coarse_target = 300;
    int max_levels = 15;
    bool fp32 = false;   // run the cycle in FP32 (coarse solve in FP64); FP32 copies made at set-up
    // The whole set-up runs on the device, level by level (D^-1, spectral radius, aggregation,
    // P smoothing, transpose, Galerkin product); only the coarse operator comes to the host for
    // its dense inverse.
    bool keep_p0_host = false;     // download the level-0 prolongator for remap() (without it, remap() sets up anew)
    bool keep_aggregates = false;  // download the level-0 aggregates into last_aggregates
    int device_levels = 0;         // levels the last set-up built
    bool force_sorted_products = false;   // every product by the global-memory sort (a check)
    int sorted_products = 0;       // products of the last set-up that took the global-memory sort
    long long front_rounds = 0;    // rounds of the aggregation's decision front in the last set-up

    std::vector<std::unique_ptr<Level>> levels;
    int coarse_n = 0;
    DevArray<double> d_coarse_inv, d_cx, d_cb;
    DevArray<float> d_cx32, d_cb32;
    double setup_ms = 0.0;
    enum { PH_DOWNLOAD, PH_RHO, PH_AGG, PH_SMOOTH, PH_TRANSPOSE, PH_GALERKIN, PH_UPLOAD, PH_COARSE, PH_REMAP, PH_COUNT };
    HostCsr P0_host;
    int remaps = 0;
    double phase_ms[PH_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> last_aggregates;   // level-0 aggregate of each row (set-up record)

    AmgPreconditioner() = default;
    AmgPreconditioner(const AmgPreconditioner&) = delete;
    AmgPreconditioner& operator=(const AmgPreconditioner&) = delete;

    int num_levels() const { return (int)levels.size() + 1; }
    void setup(int n, int nnz, const int* d_row_ptr, const int* d_col_ind, const double* d_val);
    void remap(int n, int nnz, const int* d_row_ptr, const int* d_col_ind, const double* d_val,
               const int* h_rp, const int* h_ci, const std::vector<int>& new_to_old);
    void apply(const double* d_r, double* d_z);
    const cfa_amg_t* view() const { return &view_; }
    // A view of the current hierarchy with a second set of work vectors, for a V-cycle running on
    // another stream at the same time as one through view(); valid until the next set-up or remap.
    const cfa_amg_t* second_view();

private:
    // the set-up's second stream: each level's power iteration runs on it beside the aggregation
    // (both only read the level's operator and D^-1), joined by two events
    struct AuxStream {
        cfa_stream_t s = nullptr;
        void* ready = nullptr;   // D^-1 written (recorded on the solver stream)
        void* done = nullptr;    // power iteration finished (recorded on s)
        AuxStream() = default;
        AuxStream(const AuxStream&) = delete;
        AuxStream& operator=(const AuxStream&) = delete;
        ~AuxStream();
        void ensure();
    };
    AuxStream aux_;
    struct WorkVectors {
        DevArray<double> x, b, r, Ax;
        DevArray<float> x32, b32, r32, Ax32;
    };
    std::vector<std::unique_ptr<WorkVectors>> vec2_;
    DevArray<double> d_cx2_, d_cb2_;
    DevArray<float> d_cx32_2_, d_cb32_2_;
    std::vector<cfa_amg_level_t> desc2_;
    cfa_amg_t view2_{};
    DevCsr dev_next_, dev_tmp_, dev_next2_;    // device set-up: next operator, products
    DevArray<int> d_cnt_, d_flag_, d_agg_, d_iota_, d_info_;
    DevArray<unsigned char> d_tws_, d_sws_, d_aws_, d_gws_;   // transpose, scan, aggregation, sorted product
    DevArray<unsigned char> d_sell_ws_;
    DevArray<int> d_sell_info_;
    // the sliced-ELL copy of the level's operator when its rows are at most CFA_SELL_MAX_LEN long
    void build_sell(Level& L);
    DevArray<double> d_p0v_, d_x0_, d_pw_;
    std::vector<double> x0_host_;
    int x0_dev_n_ = 0, iota_dev_n_ = 0;
    // one level: 1 built (the next operator in dev_next_, its size in n_next, nnz_next), 0 stop
    // (coarsening stalled), -1 a product beyond the device kernels' capacity
    int device_level(int n, int nnz, const int* arp, const int* aci, const double* av,
                     std::vector<std::unique_ptr<Level>>& pool, int& n_next, int& nnz_next);
    // C = A B on the device (A rows x k with a_nnz entries, B k x cols); left_max_row: the longest
    // row of A when known (rows of at most 16 entries take the merge kernel), -1 for the LDS-sort
    // kernel straight away; a row beyond both takes the global-memory sort. false: more than
    // 2^31 - 1 products.
    bool device_product(int rows, int cols, int a_nnz, const int* arp, const int* aci, const double* av,
                        const int* brp, const int* bci, const double* bv, DevCsr& C, int left_max_row);
    std::vector<std::unique_ptr<Level>> spare_;
    std::vector<cfa_amg_level_t> desc_;
    cfa_amg_t view_{};
    Clock::time_point phase_t_;
    void phase_start() { phase_t_ = Clock::now(); }
    void phase_end(int ph) {
        auto now = Clock::now();
        phase_ms[ph] += ms_between(phase_t_, now);
        phase_t_ = now;
    }
    void release();
    void refresh_view();
    static void upload_csr_from_device(int rows, int nnz, const int* rp, const int* ci, const double* v, DevCsr& D);
    static void ensure_level_vectors(Level& L, int n);
};

// lanes per row of the V-cycle kernels from the mean row length and the level size (the
// reference picks them by row length alone, src/cfa/linalg.cuh:919-926; this rule is the
// MI355X measurement's)
int csr_lanes_per_row(int rows, long long nnz);

}  // namespace cfa
