// Mesh topology on gfx950: the node-element incidence, the CSR graph, the owner-computes position
// table and the mesh checks, built from the element list after a remesh. The reference builds all
// of them on the host (src/cfa/mesh.cuh:198-376: build_csr_and_boundary_topology,
// validate_mesh); the results here are the host's exactly:
//   * incidence: a stable radix sort of the 3 ne (node, 3 e + corner) pairs by node keeps every
//     node's elements in ascending order, as the host's counting sort does;
//   * row i = {i} and the other vertices of its elements, sorted and unique (one thread per row);
//   * the checks count with integer atomics and take the minimum area and the maximum corner
//     cosine with atomics on order-preserving integer images of the doubles, so no result
//     depends on the order of the threads (the element areas' total, which is only reported, is
//     summed in a fixed tree order instead of the host's element order).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "cfa_device.hpp"

using namespace cfa_dev;

#pragma clang fp contract(off)

namespace {

constexpr int MAX_ROW = 96;   // 1 + 2 * incident elements a row may collect (47 elements)

int blocks(long long n) { return blocks_for(n > 0 ? n : 1); }
size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

unsigned key_bits(int n) {
    unsigned b = 1;
    while (b < 31 && (1u << b) < (unsigned)n) ++b;
    return b;
}

// Workspace regions (bytes, each 256-aligned): sorted node keys [3 ne], row scratch [n + 6 ne],
// row lengths [n], edge multiplicities [n + 6 ne], check cells [8 x u64], area partials [blocks],
// then the rocPRIM temporaries (the larger of the sort's and the scan's).
struct TopoWs {
    unsigned* keys;
    int* scratch;
    int* len;
    int* edge_count;
    unsigned long long* cells;
    double* area_part;
    void* tmp;
    size_t tmp_bytes;
};

size_t sort_bytes(int ne, int n) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
                                  rocprim::counting_iterator<int>(0), static_cast<int*>(nullptr), (size_t)3 * ne, 0,
                                  key_bits(n)) != hipSuccess)
        return 0;
    return bytes;
}

size_t scan_bytes(int n) {
    size_t bytes = 0;
    if (rocprim::exclusive_scan(nullptr, bytes, static_cast<const int*>(nullptr), static_cast<int*>(nullptr), 0,
                                (size_t)n + 1, rocprim::plus<int>()) != hipSuccess)
        return 0;
    return bytes;
}

size_t area_blocks(int ne) { return (size_t)blocks(ne); }

size_t ws_layout(int n, int ne, char* base, TopoWs* w) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += align_up(bytes);
        return base ? base + at : nullptr;
    };
    const size_t rows_cap = (size_t)n + 6 * (size_t)ne;
    char* keys = take(3 * (size_t)ne * sizeof(unsigned));
    char* scratch = take(rows_cap * sizeof(int));
    char* len = take(((size_t)n + 1) * sizeof(int));
    char* ec = take(rows_cap * sizeof(int));
    char* cells = take(8 * sizeof(unsigned long long));
    char* part = take(area_blocks(ne) * sizeof(double));
    const size_t tb = std::max(sort_bytes(ne, n), scan_bytes(n));
    char* tmp = take(tb);
    if (w) {
        w->keys = reinterpret_cast<unsigned*>(keys);
        w->scratch = reinterpret_cast<int*>(scratch);
        w->len = reinterpret_cast<int*>(len);
        w->edge_count = reinterpret_cast<int*>(ec);
        w->cells = reinterpret_cast<unsigned long long*>(cells);
        w->area_part = reinterpret_cast<double*>(part);
        w->tmp = tmp;
        w->tmp_bytes = tb;
    }
    return tb ? off : 0;
}

// inc_ptr[i] = first sorted position whose node is >= i, i = 0..n
__global__ void k_node_starts(int n, int total, const unsigned* __restrict__ keys, int* __restrict__ inc_ptr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    int lo = 0, hi = total;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (unsigned)i) lo = mid + 1;
        else hi = mid;
    }
    inc_ptr[i] = lo;
}

// Row i into its scratch slot (offset i + 2 inc_ptr[i], room 1 + 2 deg): {i} and the other
// vertices of its elements, insertion-sorted, duplicates dropped; len[i] = its length. A row
// that could exceed MAX_ROW entries gets length 0 and raises *overflow (the caller then builds
// the graph on the host).
__global__ void k_row_lists(int n, const int* __restrict__ ev, const int* __restrict__ inc_ptr,
                            const int* __restrict__ inc_elem, int* __restrict__ scratch, int* __restrict__ len,
                            int* __restrict__ overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {   // the scan's n + 1-th input
        len[n] = 0;
        return;
    }
    const int s0 = inc_ptr[i], s1 = inc_ptr[i + 1];
    if (1 + 2 * (s1 - s0) > MAX_ROW) {
        len[i] = 0;
        atomicOr(overflow, 1);
        return;
    }
    int row[MAX_ROW];
    int m = 0;
    row[m++] = i;
    for (int s = s0; s < s1; ++s) {
        const int e = inc_elem[s] / 3;
        for (int k = 0; k < 3; ++k) {
            const int v = ev[3 * e + k];
            if (v == i) continue;
            int p = m++;
            while (p > 0 && row[p - 1] > v) {
                row[p] = row[p - 1];
                --p;
            }
            row[p] = v;
        }
    }
    int u = 0;
    for (int k = 0; k < m; ++k)
        if (u == 0 || row[k] != row[u - 1]) row[u++] = row[k];
    int* dst = scratch + (size_t)i + 2 * (size_t)s0;
    for (int k = 0; k < u; ++k) dst[k] = row[k];
    len[i] = u;
}

__global__ void k_row_copy(int n, const int* __restrict__ inc_ptr, const int* __restrict__ scratch,
                           const int* __restrict__ row_ptr, int* __restrict__ col_ind) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int* src = scratch + (size_t)i + 2 * (size_t)inc_ptr[i];
    const int r0 = row_ptr[i], r1 = row_ptr[i + 1];
    for (int k = r0; k < r1; ++k) col_ind[k] = src[k - r0];
}

__device__ __forceinline__ int find_in_row(const int* __restrict__ ci, int lo, int hi, int col) {
    int a = lo, b = hi;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (ci[mid] < col) a = mid + 1;
        else b = mid;
    }
    return (a < hi && ci[a] == col) ? a : -1;
}

// inc_pos[3 s + b] = CSR position of column v_b in the row of incidence s's node, or -1
__global__ void k_inc_pos(int total, const int* __restrict__ ev, const unsigned* __restrict__ node_of,
                          const int* __restrict__ inc_elem, const int* __restrict__ row_ptr,
                          const int* __restrict__ col_ind, int* __restrict__ inc_pos) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= total) return;
    const int i = (int)node_of[s];
    const int e = inc_elem[s] / 3;
    const int lo = row_ptr[i], hi = row_ptr[i + 1];
    for (int b = 0; b < 3; ++b) inc_pos[3 * (size_t)s + b] = find_in_row(col_ind, lo, hi, ev[3 * e + b]);
}

// order-preserving unsigned images of doubles (for atomic min / max)
__device__ __forceinline__ unsigned long long ordered(double d) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

enum { C_NONPOS = 0, C_MISSING = 1, C_BAD = 2, C_MIN_AREA = 3, C_MAX_COS = 4 };

// Per element: signed area (non-positive ones counted), min area, area partial, the element's
// edges looked up in the CSR graph (multiplicity counted, missing ones counted), max corner cosine.
__global__ void __launch_bounds__(BLOCK)
k_check_elements(int ne, const int* __restrict__ ev, const double* __restrict__ x, const double* __restrict__ y,
                 const int* __restrict__ row_ptr, const int* __restrict__ col_ind, int* __restrict__ edge_count,
                 unsigned long long* __restrict__ cells, double* __restrict__ area_part) {
    __shared__ double lds[BLOCK / WAVE];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double area = 0.0;
    if (e < ne) {
        const int v0 = ev[3 * e], v1 = ev[3 * e + 1], v2 = ev[3 * e + 2];
        const double x0 = x[v0], y0 = y[v0], x1 = x[v1], y1 = y[v1], x2 = x[v2], y2 = y[v2];
        area = 0.5 * ((x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0));
        if (!(area > 0.0)) atomicAdd(&cells[C_NONPOS], 1ull);
        atomicMin(&cells[C_MIN_AREA], ordered(area));
        const int v[3] = {v0, v1, v2};
        double max_cos = -1.0;
        for (int k = 0; k < 3; ++k) {
            const int a = v[k], b = v[(k + 1) % 3], c = v[(k + 2) % 3];
            const int lo = min(a, b), hi = max(a, b);
            const int pos = find_in_row(col_ind, row_ptr[lo], row_ptr[lo + 1], hi);
            if (pos < 0) atomicAdd(&cells[C_MISSING], 1ull);
            else atomicAdd(&edge_count[pos], 1);
            const double ux = x[b] - x[a], uy = y[b] - y[a];
            const double wx = x[c] - x[a], wy = y[c] - y[a];
            const double cosang = (ux * wx + uy * wy) / (sqrt((ux * ux + uy * uy) * (wx * wx + wy * wy)) + 1.0e-300);
            max_cos = fmax(max_cos, cosang);
        }
        atomicMax(&cells[C_MAX_COS], ordered(max_cos));
    }
    area = block_sum(area, lds);
    if (threadIdx.x == 0) area_part[blockIdx.x] = area;
}

// Both ends on one physical boundary piece (the host rule of mesh.cpp same_boundary_piece)
__device__ bool same_boundary_piece(int ta, int tb, double xa, double ya, double xb, double yb, double H, double L) {
    const double eps = 1.0e-9;
    if (ta == BC_INTERIOR || tb == BC_INTERIOR) return false;
    if (ta == BC_CYLINDER || tb == BC_CYLINDER) return ta == BC_CYLINDER && tb == BC_CYLINDER;
    if (fabs(ya) < eps && fabs(yb) < eps) return true;
    if (fabs(ya - H) < eps && fabs(yb - H) < eps) return true;
    if (fabs(xa) < eps && fabs(xb) < eps) return true;
    if (fabs(xa - L) < eps && fabs(xb - L) < eps) return true;
    return false;
}

// Edges (a, b > a) shared by more than two elements, or by one off a boundary piece
__global__ void k_check_edges(int n, const int* __restrict__ row_ptr, const int* __restrict__ col_ind,
                              const int* __restrict__ edge_count, const int* __restrict__ tag,
                              const double* __restrict__ x, const double* __restrict__ y, double H, double L,
                              unsigned long long* __restrict__ cells) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    int bad = 0;
    for (int k = row_ptr[a]; k < row_ptr[a + 1]; ++k) {
        const int b = col_ind[k], c = edge_count[k];
        if (b <= a || c == 0) continue;
        if (c > 2 || (c == 1 && !same_boundary_piece(tag[a], tag[b], x[a], y[a], x[b], y[b], H, L))) ++bad;
    }
    if (bad) atomicAdd(&cells[C_BAD], (unsigned long long)bad);
}

__global__ void k_check_final(int nb, const double* __restrict__ part, const unsigned long long* __restrict__ cells,
                              double* __restrict__ stats) {
    __shared__ double lds[BLOCK / WAVE];
    const double total = sum_partials(part, nb, 1, lds);
    if (threadIdx.x != 0) return;
    auto back = [](unsigned long long u) {
        const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
        return __longlong_as_double((long long)b);
    };
    stats[0] = (double)cells[C_NONPOS];
    stats[1] = (double)cells[C_MISSING];
    stats[2] = (double)cells[C_BAD];
    stats[3] = back(cells[C_MIN_AREA]);
    stats[4] = total;
    stats[5] = back(cells[C_MAX_COS]);
}

__global__ void k_zero_int(int* __restrict__ p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = 0;
}

// *nnz_out = the graph's entries, or -1 when a row did not fit
__global__ void k_topology_done(int n, const int* __restrict__ row_ptr, const int* __restrict__ overflow,
                                int* __restrict__ nnz_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) nnz_out[0] = overflow[0] ? -1 : row_ptr[n];
}

__global__ void k_check_init(int count, int* __restrict__ edge_count, unsigned long long* __restrict__ cells) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) edge_count[k] = 0;
    if (k == 0) {
        cells[C_NONPOS] = 0;
        cells[C_MISSING] = 0;
        cells[C_BAD] = 0;
        cells[C_MIN_AREA] = 0xffffffffffffffffull;
        cells[C_MAX_COS] = 0;
    }
}

}  // namespace

extern "C" {

size_t cfa_topology_ws_bytes(int n_nodes, int n_elems) {
    if (n_nodes <= 0 || n_elems <= 0) return 0;
    return ws_layout(n_nodes, n_elems, nullptr, nullptr);
}

int cfa_mesh_topology(int n, int ne, const int* elem_v, void* ws, size_t ws_bytes, int* inc_ptr, int* inc_elem,
                      int* row_ptr, int* col_ind, int* inc_pos, int* nnz_out, cfa_stream_t stream) {
    if (n <= 0 || ne <= 0 || !elem_v || !ws || !inc_ptr || !inc_elem || !row_ptr || !col_ind || !inc_pos || !nnz_out)
        return CFA_E_ARG;
    const size_t need = cfa_topology_ws_bytes(n, ne);
    if (need == 0 || ws_bytes < need) return CFA_E_ARG;
    TopoWs w;
    ws_layout(n, ne, static_cast<char*>(ws), &w);
    const hipStream_t s = as_stream(stream);
    const int total = 3 * ne;
    size_t tb = w.tmp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(w.tmp, tb, reinterpret_cast<const unsigned*>(elem_v), w.keys,
                                             rocprim::counting_iterator<int>(0), inc_elem, (size_t)total, 0,
                                             key_bits(n), s);
    if (e != hipSuccess) return (int)e;
    int* overflow = reinterpret_cast<int*>(w.cells + 7);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(WAVE), 0, s, overflow);
    hipLaunchKernelGGL(k_node_starts, dim3(blocks(n + 1)), dim3(BLOCK), 0, s, n, total, w.keys, inc_ptr);
    hipLaunchKernelGGL(k_row_lists, dim3(blocks(n + 1)), dim3(BLOCK), 0, s, n, elem_v, inc_ptr, inc_elem, w.scratch,
                       w.len, overflow);
    tb = w.tmp_bytes;
    e = rocprim::exclusive_scan(w.tmp, tb, w.len, row_ptr, 0, (size_t)n + 1, rocprim::plus<int>(), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_row_copy, dim3(blocks(n)), dim3(BLOCK), 0, s, n, inc_ptr, w.scratch, row_ptr, col_ind);
    hipLaunchKernelGGL(k_inc_pos, dim3(blocks(total)), dim3(BLOCK), 0, s, total, elem_v, w.keys, inc_elem, row_ptr,
                       col_ind, inc_pos);
    hipLaunchKernelGGL(k_topology_done, dim3(1), dim3(WAVE), 0, s, n, row_ptr, overflow, nnz_out);
    return launch_status();
}

int cfa_mesh_tables(int n, int ne, const int* elem_v, const int* row_ptr, const int* col_ind, void* ws,
                    size_t ws_bytes, int* inc_ptr, int* inc_elem, int* inc_pos, cfa_stream_t stream) {
    if (n <= 0 || ne <= 0 || !elem_v || !row_ptr || !col_ind || !ws || !inc_ptr || !inc_elem || !inc_pos)
        return CFA_E_ARG;
    const size_t need = cfa_topology_ws_bytes(n, ne);
    if (need == 0 || ws_bytes < need) return CFA_E_ARG;
    TopoWs w;
    ws_layout(n, ne, static_cast<char*>(ws), &w);
    const hipStream_t s = as_stream(stream);
    const int total = 3 * ne;
    size_t tb = w.tmp_bytes;
    const hipError_t e = rocprim::radix_sort_pairs(w.tmp, tb, reinterpret_cast<const unsigned*>(elem_v), w.keys,
                                                   rocprim::counting_iterator<int>(0), inc_elem, (size_t)total, 0,
                                                   key_bits(n), s);
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_node_starts, dim3(blocks(n + 1)), dim3(BLOCK), 0, s, n, total, w.keys, inc_ptr);
    hipLaunchKernelGGL(k_inc_pos, dim3(blocks(total)), dim3(BLOCK), 0, s, total, elem_v, w.keys, inc_elem, row_ptr,
                       col_ind, inc_pos);
    return launch_status();
}

int cfa_mesh_check(int n, int ne, const int* elem_v, const double* x, const double* y, const int* bc_tag,
                   const int* row_ptr, const int* col_ind, int nnz, double h_channel, double l_channel, void* ws,
                   size_t ws_bytes, double* stats, cfa_stream_t stream) {
    if (n <= 0 || ne <= 0 || nnz < 0 || !elem_v || !x || !y || !bc_tag || !row_ptr || !col_ind || !ws || !stats)
        return CFA_E_ARG;
    const size_t need = cfa_topology_ws_bytes(n, ne);
    if (need == 0 || ws_bytes < need || (size_t)nnz > (size_t)n + 6 * (size_t)ne) return CFA_E_ARG;
    TopoWs w;
    ws_layout(n, ne, static_cast<char*>(ws), &w);
    const hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_check_init, dim3(blocks(nnz > 0 ? nnz : 1)), dim3(BLOCK), 0, s, nnz, w.edge_count, w.cells);
    const int nb = (int)area_blocks(ne);
    hipLaunchKernelGGL(k_check_elements, dim3(nb), dim3(BLOCK), 0, s, ne, elem_v, x, y, row_ptr, col_ind,
                       w.edge_count, w.cells, w.area_part);
    hipLaunchKernelGGL(k_check_edges, dim3(blocks(n)), dim3(BLOCK), 0, s, n, row_ptr, col_ind, w.edge_count, bc_tag,
                       x, y, h_channel, l_channel, w.cells);
    hipLaunchKernelGGL(k_check_final, dim3(1), dim3(BLOCK), 0, s, nb, w.area_part, w.cells, stats);
    return launch_status();
}

size_t cfa_sort_pairs_ws_bytes(int n) {
    if (n <= 0) return 0;
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const unsigned*>(nullptr), static_cast<unsigned*>(nullptr),
                                  static_cast<const int*>(nullptr), static_cast<int*>(nullptr), (size_t)n) != hipSuccess)
        return 0;
    return bytes;
}

int cfa_sort_pairs_u32(int n, const unsigned* keys_in, const int* vals_in, unsigned* keys_out, int* vals_out,
                       void* ws, size_t ws_bytes, cfa_stream_t stream) {
    if (n < 0 || (n > 0 && (!keys_in || !vals_in || !keys_out || !vals_out || !ws))) return CFA_E_ARG;
    if (n == 0) return CFA_OK;
    size_t need = cfa_sort_pairs_ws_bytes(n);
    if (need == 0 || ws_bytes < need) return CFA_E_ARG;
    const hipError_t e =
        rocprim::radix_sort_pairs(ws, need, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0, 32, as_stream(stream));
    return e == hipSuccess ? launch_status() : (int)e;
}

}  // extern "C"
