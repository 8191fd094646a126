// Newest-vertex-bisection refinement forest with coarsening, advective look-ahead, curved
// cylinder midpoints, P1 solution transfer and Morton-ordered extraction.
//
// Interface of the reference's host AMR layer (src/cfa/amr.cuh:9-471): AmrParams, AmrStats,
// AdaptiveMesh(base, u0, v0, p0), adapt, refine_uniform, extract, scatter_solution,
// retransfer, node_vertices, num_vertex_slots, inlet_umax, dbg_ms. Every integer decision is
// made with the same rules and in the same order, so meshes and their numbering match the
// reference bit for bit for the same indicators.
#pragma once

#include <array>
#include <atomic>
#include <functional>
#include <mutex>
#include <unordered_map>

#include "mesh.hpp"

namespace cfa {

struct AmrParams {
    int max_level = 4;
    int min_level = 2;
    double tol_vort = 0.1;
    double tol_interp = 0.01;
    double coarsen_safety = 0.35;
    int buffer_layers = 1;
    int max_elems = 200000;
    double lookahead_time = 0.0;
    double keep_radius = 0.25;
    double tol_radius = 0.0;
    double tol_power = 1.0;
};

struct AmrStats {
    int leaves_before = 0;
    int leaves_after = 0;
    int marked_refine = 0;
    int marked_coarsen = 0;
    int bisections = 0;
    int closure_bisections = 0;
    int removed_vertices = 0;
    double elapsed_ms = 0.0;
    double ms_lookahead = 0.0;
    double ms_mark = 0.0;
    double ms_neighbours = 0.0;
    double ms_coarsen = 0.0;
    double ms_refine = 0.0;
};

struct AmrElem {
    int v[3] = {-1, -1, -1};
    int gen = 0;
    int parent = -1;
    int child[2] = {-1, -1};
    bool alive = false;
    bool is_leaf() const { return alive && child[0] < 0; }
};

class AdaptiveMesh {
public:
    enum { D_ADJ, D_TRACE, D_MERGE, D_COLLECT, D_KEYS, D_SORT, D_COMPACT, D_FILL, D_COUNT };
    double dbg_ms[D_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    double inlet_umax = Sim::U_MAX;   // peak inflow imposed on new inlet midpoints

    AdaptiveMesh(const Mesh2D& base, const std::vector<double>& u0, const std::vector<double>& v0,
                 const std::vector<double>& p0);

    int num_leaves() const { return (int)leaf_ids_.size(); }
    const std::vector<int>& node_vertices() const { return vert_ids_; }
    int num_vertex_slots() const { return (int)x_.size(); }

    void scatter_solution(const std::vector<double>& h_u, const std::vector<double>& h_v,
                          const std::vector<double>& h_p, const std::vector<double>& h_uo,
                          const std::vector<double>& h_vo);
    void retransfer(const std::vector<int>& old_vids, const std::vector<double>& h_u,
                    const std::vector<double>& h_v, const std::vector<double>& h_p,
                    const std::vector<double>& h_uo, const std::vector<double>& h_vo,
                    std::vector<double>& n_u, std::vector<double>& n_v, std::vector<double>& n_p,
                    std::vector<double>& n_uo, std::vector<double>& n_vo);
    // Sorts extract's (Morton key, forest id) pairs ascending in place (ids arrive ascending, so a
    // stable sort by key is the same order); when unset or returning false, extract sorts on the
    // host. The driver sets a device sorter for synchronous remeshes only: with overlap the
    // extraction runs on the worker thread, which does not use the GPU.
    std::function<bool(std::vector<std::pair<uint32_t, int>>&)> leaf_sorter;
    void extract(Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                 std::vector<double>& h_p, std::vector<double>& h_uo, std::vector<double>& h_vo);
    AmrStats adapt(const std::vector<double>& eta, const AmrParams& prm);
    AmrStats refine_uniform(int levels);

private:
    // persistent vertex slots (recycled after coarsening)
    std::vector<double> x_, y_, u_, v_, p_, uo_, vo_;
    std::vector<int> bc_tag_;
    std::vector<char> v_alive_;
    std::vector<std::array<int, 2>> v_parents_;   // elements bisected through this vertex
    std::vector<uint64_t> v_edge_;                 // bisected edge of a midpoint vertex
    std::vector<int> free_verts_;
    // refinement forest (recycled slots)
    std::vector<AmrElem> elems_;
    std::vector<int> free_elems_;
    std::unordered_map<uint64_t, int> edge_mid_;
    std::vector<std::array<int, 3>> created_log_;   // {m, a, b} per midpoint of the last adapt
    // compact numbering of the last extract()
    std::vector<int> leaf_ids_, vert_ids_;
    // leaf adjacency of the current adapt()
    std::vector<int> leaf_pos_, leaf_nbr_, inc_ptr_, inc_list_;
    // vertex -> leaves at the start of refinement
    std::vector<int> vleaf_ptr_, vleaf_list_;
    std::vector<int> search_stack_;
    // bucket grid (look-ahead fallback), built on first use within an adapt()
    static constexpr double BUCKET_CELL = 0.01;
    std::atomic<bool> buckets_built_{false};
    std::mutex buckets_mu_;
    int bucket_nx_ = 0, bucket_ny_ = 0;
    std::vector<int> bucket_ptr_, bucket_items_;
    Clock::time_point dbg_t_;

    void dbg_start() { dbg_t_ = Clock::now(); }
    void dbg_lap(int k) {
        auto now = Clock::now();
        dbg_ms[k] += ms_between(dbg_t_, now);
        dbg_t_ = now;
    }

    static uint32_t spread_bits(uint32_t n);
    static uint32_t morton_code(double sx, double sy);
    double leaf_area(int e) const;
    void build_leaf_adjacency();
    bool barycentric(int e, double px, double py, double l[3]) const;
    void build_buckets();
    int locate_bucket(double px, double py);
    int locate_from(int start, double px, double py);
    void apply_lookahead(std::vector<double>* av, std::vector<double>* ai, double horizon,
                         double q_v, double q_i);
    int alloc_vertex();
    int alloc_elem();
    void free_elem(int e);
    int get_or_create_midpoint(int a, int b);
    void set_midpoint_values(int m, int a, int b);
    void bisect(int e, std::vector<int>* new_children, std::vector<std::pair<int, int>>* new_edges);
    void build_vertex_leaves();
    bool has_vertex(int e, int a) const {
        const int* w = elems_[e].v;
        return w[0] == a || w[1] == a || w[2] == a;
    }
    void leaves_with_edge(int a, int b, std::vector<int>& out);
    bool is_hanging(int e) const;
    int make_conforming(std::vector<int>& children, std::vector<std::pair<int, int>>& edges);
    void collapse_vertex(int m, std::vector<char>& coarsen, std::vector<double>& av,
                         std::vector<double>& ai);
    void refine_to_targets(std::vector<int>& target, AmrStats& st);
};

}  // namespace cfa
