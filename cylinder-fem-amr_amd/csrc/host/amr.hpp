// Adaptive mesh: newest-vertex-bisection refinement forest with coarsening, advective
// look-ahead, curved cylinder midpoints, P1 solution transfer and Morton-ordered extraction.
//
// Interface of the reference's host AMR layer (src/cfa/amr.cuh:9-471): AmrParams, AmrStats,
// AdaptiveMesh(base, u0, v0, p0), adapt, refine_uniform, extract, scatter_solution, retransfer,
// node_vertices, num_vertex_slots, inlet_umax, dbg_ms. The forest lives on the device and every
// phase runs as gfx950 kernels (cfa_nvb_* in include/cfa_hip.h) on the object's own stream, which
// does not synchronise with the solver's: an adaptation may run on a worker thread while the
// solver keeps stepping. Meshes, numbering and slot recycling follow the reference's rules.
#pragma once

#include <memory>

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

class AdaptiveMesh {
public:
    // accumulated device phase times [ms]: marking (indicators, adjacency, look-ahead, marks,
    // coarsening flags), coarsening sweeps, refinement passes, leaf lists, extraction, download
    enum { D_MARK, D_COARSEN, D_REFINE, D_LEAVES, D_EXTRACT, D_DOWNLOAD, D_COUNT };
    double dbg_ms[D_COUNT] = {0, 0, 0, 0, 0, 0};
    double inlet_umax = Sim::U_MAX;   // peak inflow imposed on new inlet midpoints

    AdaptiveMesh(const Mesh2D& base, const std::vector<double>& u0, const std::vector<double>& v0,
                 const std::vector<double>& p0);
    ~AdaptiveMesh();
    AdaptiveMesh(const AdaptiveMesh&) = delete;
    AdaptiveMesh& operator=(const AdaptiveMesh&) = delete;

    int num_leaves() const;
    // compact node of the last extract() -> persistent vertex slot
    const std::vector<int>& node_vertices() const { return vert_ids_; }
    // compact element of the last extract() -> forest element slot
    const std::vector<int>& leaf_elements() const { return leaf_ids_; }
    int num_vertex_slots() const;
    // slots in use and recycled slots waiting for reuse: {elements, vertices, free elements, free vertices}
    void slot_counts(int out[4]) const;
    // bisections of each refinement pass of the last adapt() / refine_uniform(), in order
    const std::vector<int>& pass_sizes() const;

    void scatter_solution(const std::vector<double>& h_u, const std::vector<double>& h_v,
                          const std::vector<double>& h_p, const std::vector<double>& h_uo,
                          const std::vector<double>& h_vo);
    void retransfer(const std::vector<int>& old_vids, const std::vector<double>& h_u,
                    const std::vector<double>& h_v, const std::vector<double>& h_p,
                    const std::vector<double>& h_uo, const std::vector<double>& h_vo,
                    std::vector<double>& n_u, std::vector<double>& n_v, std::vector<double>& n_p,
                    std::vector<double>& n_uo, std::vector<double>& n_vo);
    void extract(Mesh2D& mesh, std::vector<double>& h_u, std::vector<double>& h_v,
                 std::vector<double>& h_p, std::vector<double>& h_uo, std::vector<double>& h_vo);
    AmrStats adapt(const std::vector<double>& eta, const AmrParams& prm);
    AmrStats refine_uniform(int levels);

private:
    struct Device;
    std::unique_ptr<Device> d_;
    std::vector<int> vert_ids_, leaf_ids_;
};

}  // namespace cfa
