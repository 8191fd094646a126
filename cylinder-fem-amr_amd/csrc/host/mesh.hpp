// Compact solver mesh, base-mesh generator, node-adjacency CSR graph and mesh validation.
//
// Interface kept from the reference so a caller of its host mesh layer finds the same names and
// meanings: Triangle / Mesh2D (src/cfa/mesh.cuh:8-41), generate_cylinder_mesh (:83-193),
// build_csr_and_boundary_topology (:198-270), MeshQuality / validate_mesh (:275-376).
#pragma once

#include "cfa_common.hpp"

namespace cfa {

struct Triangle {
    int v[3];     // counter-clockwise; v[0] newest vertex, (v[1], v[2]) refinement edge
    int level;    // gen / 2 (one level halves h)
    int gen = 0;  // bisections since the base-mesh ancestor
};

struct Mesh2D {
    int num_nodes = 0;
    int num_elems = 0;
    std::vector<double> x, y;
    std::vector<int> bc_tag;
    std::vector<Triangle> elems;
    // node-to-node adjacency (row i: i and its edge neighbours, ascending)
    int nnz = 0;
    std::vector<int> csr_row_ptr;
    std::vector<int> csr_col_ind;
    // cylinder boundary edges (a < b) and the element owning each
    std::vector<std::pair<int, int>> cylinder_edges;
    std::vector<int> cylinder_edge_elem;
};

struct MeshQuality {
    double min_angle_deg = 180.0;
    double min_area = 1.0e300;
    double total_area = 0.0;
    int max_level = 0;
    int max_gen = 0;
};

double signed_area(const Mesh2D& mesh, const Triangle& tri);
int classify_boundary_node(double px, double py);

// Body-fitted O-grid block around the cylinder plus a structured wake block; every quad is cut
// along one diagonal that becomes the refinement edge of both halves (matching NVB labels).
void generate_cylinder_mesh(Mesh2D& mesh, int n_quarter = 8, int n_radial = 7, int n_wake_x = 32);

// Sorted-row node adjacency and the cylinder boundary edges (edges of exactly one element).
void build_csr_and_boundary_topology(Mesh2D& mesh);
// the cylinder edge list alone (the part of build_csr_and_boundary_topology after the graph)
void build_boundary_topology(Mesh2D& mesh);

// True for a conforming, positively oriented triangulation; fills q with quality figures.
bool validate_mesh(const Mesh2D& mesh, MeshQuality& q);

// build_csr_and_boundary_topology + validate_mesh, aborting on an invalid mesh
// (src/cfa/device.cuh:491-499).
MeshQuality finalize_mesh(Mesh2D& mesh);

}  // namespace cfa
