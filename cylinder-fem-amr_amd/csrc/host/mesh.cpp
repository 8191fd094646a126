// Base mesh, CSR graph and validation. The node coordinates, the node and element order and the
// refinement-edge labels must come out bit-identical to src/cfa/mesh.cuh:83-193, because every
// later integer decision (NVB, Morton order, colouring, aggregation) and the reference's stored
// force histories depend on them; the floating-point expressions are therefore evaluated in the
// same form and order.
#include "mesh.hpp"

#include <algorithm>
#include <unordered_map>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cfa {

double signed_area(const Mesh2D& mesh, const Triangle& tri) {
    const double x0 = mesh.x[tri.v[0]], y0 = mesh.y[tri.v[0]];
    const double x1 = mesh.x[tri.v[1]], y1 = mesh.y[tri.v[1]];
    const double x2 = mesh.x[tri.v[2]], y2 = mesh.y[tri.v[2]];
    return 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
}

// src/cfa/mesh.cuh:62-70: cylinder by radius (1e-4), then walls, inlet, outlet by coordinate.
int classify_boundary_node(double px, double py) {
    const double eps = 1.0e-6;
    const double r = std::hypot(px - Sim::X_CYL, py - Sim::Y_CYL);
    if (std::abs(r - Sim::R_CYL) < 1.0e-4) return BC_CYLINDER;
    if (std::abs(py) < eps || std::abs(py - Sim::H_CHANNEL) < eps) return BC_WALL;
    if (std::abs(px) < eps) return BC_INLET;
    if (std::abs(px - Sim::L_CHANNEL) < eps) return BC_OUTLET;
    return BC_INTERIOR;
}

namespace {

struct MeshBuilder {
    Mesh2D& m;
    // Node insertion with de-duplication: an existing node within 1e-7 (first in insertion order)
    // is reused, as in src/cfa/mesh.cuh:93-106.
    int node(double px, double py, int tag_override = -1) {
        for (int i = 0; i < (int)m.x.size(); ++i)
            if (std::hypot(m.x[i] - px, m.y[i] - py) < 1.0e-7) return i;
        m.x.push_back(px);
        m.y.push_back(py);
        m.bc_tag.push_back(tag_override >= 0 ? tag_override : classify_boundary_node(px, py));
        return (int)m.x.size() - 1;
    }
    // Quad (p00, p10, p11, p01) cut along p00-p11; both halves take the diagonal as their
    // refinement edge (newest vertex opposite), src/cfa/mesh.cuh:109-112.
    void quad(int p00, int p10, int p11, int p01) {
        Triangle a;  a.v[0] = p10;  a.v[1] = p11;  a.v[2] = p00;  a.level = 0;  a.gen = 0;
        Triangle b;  b.v[0] = p01;  b.v[1] = p00;  b.v[2] = p11;  b.level = 0;  b.gen = 0;
        m.elems.push_back(a);
        m.elems.push_back(b);
    }
};

}  // namespace

void generate_cylinder_mesh(Mesh2D& mesh, int n_quarter, int n_radial, int n_wake_x) {
    mesh = Mesh2D();
    MeshBuilder mb{mesh};
    const double box = Sim::H_CHANNEL;   // square block [0, H] x [0, H] around the cylinder
    const int n_azi = 4 * n_quarter;

    // Outer square, counter-clockwise from the south-east corner: east, north, west, south side.
    std::vector<double> bx(n_azi), by(n_azi), cx(n_azi), cy(n_azi);
    for (int q = 0; q < n_quarter; ++q) {
        const double s = (double)q / n_quarter;
        bx[q] = box;                          by[q] = s * box;
        bx[n_quarter + q] = (1.0 - s) * box;  by[n_quarter + q] = box;
        bx[2 * n_quarter + q] = 0.0;          by[2 * n_quarter + q] = (1.0 - s) * box;
        bx[3 * n_quarter + q] = s * box;      by[3 * n_quarter + q] = 0.0;
    }
    // Cylinder points on the rays from the centre to the square's points.
    for (int i = 0; i < n_azi; ++i) {
        const double th = std::atan2(by[i] - Sim::Y_CYL, bx[i] - Sim::X_CYL);
        cx[i] = Sim::X_CYL + Sim::R_CYL * std::cos(th);
        cy[i] = Sim::Y_CYL + Sim::R_CYL * std::sin(th);
    }
    // Rings r = 0 (cylinder) .. n_radial (square), graded towards the wall.
    std::vector<int> ring((size_t)(n_radial + 1) * n_azi);
    for (int r = 0; r <= n_radial; ++r) {
        const double xi = (double)r / n_radial;
        const double w = 0.65 * xi + 0.35 * xi * xi;
        for (int i = 0; i < n_azi; ++i) {
            const double px = (1.0 - w) * cx[i] + w * bx[i];
            const double py = (1.0 - w) * cy[i] + w * by[i];
            ring[(size_t)r * n_azi + i] = mb.node(px, py, r == 0 ? BC_CYLINDER : -1);
        }
    }
    for (int r = 0; r < n_radial; ++r)
        for (int i = 0; i < n_azi; ++i) {
            const int j = (i + 1) % n_azi;
            mb.quad(ring[(size_t)r * n_azi + i], ring[(size_t)(r + 1) * n_azi + i],
                    ring[(size_t)(r + 1) * n_azi + j], ring[(size_t)r * n_azi + j]);
        }

    // Wake block [H, L] x [0, H]: n_quarter + 1 points per column (they match the east side),
    // columns stretched downstream.
    const int ny = n_quarter;
    std::vector<int> wake((size_t)(n_wake_x + 1) * (ny + 1));
    for (int ix = 0; ix <= n_wake_x; ++ix) {
        const double s = (double)ix / n_wake_x;
        const double xc = box + (Sim::L_CHANNEL - box) * (0.75 * s + 0.25 * s * s);
        for (int iy = 0; iy <= ny; ++iy) {
            const double yc = ((double)iy / ny) * Sim::H_CHANNEL;
            wake[(size_t)ix * (ny + 1) + iy] = mb.node(xc, yc);
        }
    }
    for (int ix = 0; ix < n_wake_x; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            mb.quad(wake[(size_t)ix * (ny + 1) + iy], wake[(size_t)(ix + 1) * (ny + 1) + iy],
                    wake[(size_t)(ix + 1) * (ny + 1) + iy + 1], wake[(size_t)ix * (ny + 1) + iy + 1]);

    mesh.num_nodes = (int)mesh.x.size();
    mesh.num_elems = (int)mesh.elems.size();
    // Orientation fix that keeps v[0] and the refinement edge: swap v[1] and v[2].
    for (Triangle& t : mesh.elems)
        if (signed_area(mesh, t) < 0.0) std::swap(t.v[1], t.v[2]);
}

void build_csr_and_boundary_topology(Mesh2D& mesh) {
    mesh.num_nodes = (int)mesh.x.size();
    mesh.num_elems = (int)mesh.elems.size();
    const int n = mesh.num_nodes, ne = mesh.num_elems;

    // node -> incident elements (counting sort by node)
    std::vector<int> inc_ptr(n + 1, 0), inc((size_t)3 * ne);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) inc_ptr[mesh.elems[e].v[k] + 1]++;
    for (int i = 0; i < n; ++i) inc_ptr[i + 1] += inc_ptr[i];
    {
        std::vector<int> pos(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) inc[pos[mesh.elems[e].v[k]]++] = e;
    }

    // Row i = {i} U vertices of its elements, sorted and unique; built in a scratch slot of
    // worst-case size 1 + 2 * (incident elements), then packed.
    std::vector<int> slot(n + 1, 0), len(n);
    for (int i = 0; i < n; ++i) slot[i + 1] = slot[i] + 1 + 2 * (inc_ptr[i + 1] - inc_ptr[i]);
    std::vector<int> scratch(slot[n]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        int* row = scratch.data() + slot[i];
        int m = 0;
        row[m++] = i;
        for (int s = inc_ptr[i]; s < inc_ptr[i + 1]; ++s) {
            const Triangle& t = mesh.elems[inc[s]];
            for (int k = 0; k < 3; ++k)
                if (t.v[k] != i) row[m++] = t.v[k];
        }
        std::sort(row, row + m);
        len[i] = (int)(std::unique(row, row + m) - row);
    }
    mesh.csr_row_ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) mesh.csr_row_ptr[i + 1] = mesh.csr_row_ptr[i] + len[i];
    mesh.csr_col_ind.resize(mesh.csr_row_ptr[n]);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        std::copy(scratch.begin() + slot[i], scratch.begin() + slot[i] + len[i],
                  mesh.csr_col_ind.begin() + mesh.csr_row_ptr[i]);
    mesh.nnz = (int)mesh.csr_col_ind.size();
    build_boundary_topology(mesh);
}

void build_boundary_topology(Mesh2D& mesh) {
    const int ne = (int)mesh.elems.size();
    // Cylinder edges: element edges whose two ends are cylinder nodes and that belong to exactly
    // one element, in order of first appearance (src/cfa/mesh.cuh:251-269).
    std::unordered_map<uint64_t, std::pair<int, int>> seen;   // key -> (count, last element)
    std::vector<uint64_t> first_order;
    for (int e = 0; e < ne; ++e) {
        const Triangle& t = mesh.elems[e];
        for (int a = 0; a < 3; ++a) {
            const int p = t.v[a], q = t.v[(a + 1) % 3];
            if (mesh.bc_tag[p] != BC_CYLINDER || mesh.bc_tag[q] != BC_CYLINDER) continue;
            const uint64_t key = edge_key(p, q);
            auto it = seen.find(key);
            if (it == seen.end()) {
                seen.emplace(key, std::make_pair(1, e));
                first_order.push_back(key);
            } else {
                it->second.first++;
                it->second.second = e;
            }
        }
    }
    mesh.cylinder_edges.clear();
    mesh.cylinder_edge_elem.clear();
    for (uint64_t key : first_order) {
        const auto& ce = seen[key];
        if (ce.first != 1) continue;
        mesh.cylinder_edges.push_back({(int)(key >> 32), (int)(key & 0xffffffffu)});
        mesh.cylinder_edge_elem.push_back(ce.second);
    }
}

namespace {
// Both ends on one physical boundary piece (src/cfa/mesh.cuh:283-294).
bool same_boundary_piece(const Mesh2D& m, int a, int b) {
    const double eps = 1.0e-9;
    const int ta = m.bc_tag[a], tb = m.bc_tag[b];
    if (ta == BC_INTERIOR || tb == BC_INTERIOR) return false;
    if (ta == BC_CYLINDER || tb == BC_CYLINDER) return ta == BC_CYLINDER && tb == BC_CYLINDER;
    const double xa = m.x[a], ya = m.y[a], xb = m.x[b], yb = m.y[b];
    if (std::abs(ya) < eps && std::abs(yb) < eps) return true;
    if (std::abs(ya - Sim::H_CHANNEL) < eps && std::abs(yb - Sim::H_CHANNEL) < eps) return true;
    if (std::abs(xa) < eps && std::abs(xb) < eps) return true;
    if (std::abs(xa - Sim::L_CHANNEL) < eps && std::abs(xb - Sim::L_CHANNEL) < eps) return true;
    return false;
}
}  // namespace

bool validate_mesh(const Mesh2D& mesh, MeshQuality& q) {
    q = MeshQuality();
    if ((int)mesh.csr_row_ptr.size() != mesh.num_nodes + 1) {
        std::fprintf(stderr, "[Mesh] CSR graph does not match the mesh\n");
        return false;
    }
    bool ok = true;
    std::vector<int> edge_count(mesh.nnz, 0);   // multiplicity of edge (a < b) at CSR (a, b)
    double max_cos = -1.0;
    for (int e = 0; e < mesh.num_elems; ++e) {
        const Triangle& t = mesh.elems[e];
        const double area = signed_area(mesh, t);
        if (!(area > 0.0)) {
            std::fprintf(stderr, "[Mesh] element %d has non-positive area %.3e\n", e, area);
            ok = false;
        }
        q.min_area = std::min(q.min_area, area);
        q.total_area += area;
        q.max_level = std::max(q.max_level, t.level);
        q.max_gen = std::max(q.max_gen, t.gen);
        for (int k = 0; k < 3; ++k) {
            const int a = t.v[k], b = t.v[(k + 1) % 3], c = t.v[(k + 2) % 3];
            const int lo = std::min(a, b), hi = std::max(a, b);
            const int* first = mesh.csr_col_ind.data() + mesh.csr_row_ptr[lo];
            const int* last = mesh.csr_col_ind.data() + mesh.csr_row_ptr[lo + 1];
            const int* pos = std::lower_bound(first, last, hi);
            if (pos == last || *pos != hi) {
                std::fprintf(stderr, "[Mesh] edge (%d, %d) missing from the CSR graph\n", lo, hi);
                ok = false;
            } else {
                edge_count[pos - mesh.csr_col_ind.data()]++;
            }
            const double ux = mesh.x[b] - mesh.x[a], uy = mesh.y[b] - mesh.y[a];
            const double wx = mesh.x[c] - mesh.x[a], wy = mesh.y[c] - mesh.y[a];
            const double cosang = (ux * wx + uy * wy) /
                                  (std::sqrt((ux * ux + uy * uy) * (wx * wx + wy * wy)) + 1.0e-300);
            max_cos = std::max(max_cos, cosang);
        }
    }
    q.min_angle_deg = std::acos(std::max(-1.0, std::min(1.0, max_cos))) * 180.0 / Sim::PI;

    int bad = 0;
    for (int a = 0; a < mesh.num_nodes; ++a)
        for (int idx = mesh.csr_row_ptr[a]; idx < mesh.csr_row_ptr[a + 1]; ++idx) {
            const int b = mesh.csr_col_ind[idx], count = edge_count[idx];
            if (b <= a || count == 0) continue;
            if (count > 2 || (count == 1 && !same_boundary_piece(mesh, a, b))) {
                if (bad < 10)
                    std::fprintf(stderr, "[Mesh] non-conforming edge (%d, %d) shared by %d element(s)\n",
                                 a, b, count);
                ++bad;
            }
        }
    if (bad > 0) {
        std::fprintf(stderr, "[Mesh] %d non-conforming edges\n", bad);
        ok = false;
    }
    return ok;
}

MeshQuality finalize_mesh(Mesh2D& mesh) {
    build_csr_and_boundary_topology(mesh);
    MeshQuality q;
    if (!validate_mesh(mesh, q)) fatal("[AMR] mesh validation failed, aborting");
    return q;
}

}  // namespace cfa
