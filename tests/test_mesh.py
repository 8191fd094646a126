"""Mesh layer against the reference's known answers, and mesh invariants checked by brute force: the
generator and CSR graph of the product's host layer with the oracle's restatement of the NVB rules
(uniform refinement, Morton extraction) on the CPU, and the product's device NVB forest against
the oracle's meshes on the GPU (connectivity, tags, graph and coordinates identical).

Known answers: base mesh 508 nodes / 896 triangles / 3316 CSR nonzeros, min angle 15.29 deg
(src/cylinder_fem_amr.cu init line in every runs/*/run.log); node counts 1912, 7408, 29152,
115648, 460672 for uniform levels 1-5 with 896 * 4^L triangles; h_max per level from the
reference's runs/paper/mms_L*/cylinder_fem_amr_mms.csv.
"""
import numpy as np
import pytest

import cfa
import cfa_oracle
from conftest import read_csv

NODES = {0: 508, 1: 1912, 2: 7408, 3: 29152, 4: 115648}
NNZ = {0: 3316, 1: 12904, 2: 50896, 3: 202144, 4: 805696}


@pytest.fixture(scope="module", params=[0, 1, 2, 3])
def mesh(request):
    return request.param, cfa_oracle.uniform_mesh(request.param)


def test_sizes(mesh):
    level, m = mesh
    assert m["n"] == NODES[level]
    assert m["ne"] == 896 * 4 ** level
    assert m["nnz"] == NNZ[level]
    if level == 0:
        assert round(m["min_angle_deg"], 2) == 15.29


def test_h_max_matches_reference(golden):
    for level in (1, 2, 3):
        m = cfa_oracle.uniform_mesh(level)
        ev = m["ev"].reshape(-1, 3)
        h = 0.0
        for a, b in ((0, 1), (1, 2), (2, 0)):
            h = max(h, np.hypot(m["x"][ev[:, a]] - m["x"][ev[:, b]], m["y"][ev[:, a]] - m["y"][ev[:, b]]).max())
        ref = float(read_csv(golden / f"mms_L{level}_mms.csv")["h_max"])
        assert abs(h - ref) <= 1e-9 * ref


def test_conforming_positive_and_area(mesh):
    """Each edge is shared by at most two triangles, boundary edges lie on the channel walls or the
    cylinder, all triangles are counter-clockwise and the areas sum to the domain's area."""
    _, m = mesh
    ev = m["ev"].reshape(-1, 3)
    x, y = m["x"], m["y"]
    area = 0.5 * ((x[ev[:, 1]] - x[ev[:, 0]]) * (y[ev[:, 2]] - y[ev[:, 0]])
                  - (x[ev[:, 2]] - x[ev[:, 0]]) * (y[ev[:, 1]] - y[ev[:, 0]]))
    assert (area > 0).all()
    edges = np.sort(np.concatenate([ev[:, [0, 1]], ev[:, [1, 2]], ev[:, [2, 0]]]), axis=1)
    uniq, count = np.unique(edges, axis=0, return_counts=True)
    assert count.max() <= 2
    bnd = uniq[count == 1]
    mx, my = 0.5 * (x[bnd[:, 0]] + x[bnd[:, 1]]), 0.5 * (y[bnd[:, 0]] + y[bnd[:, 1]])
    on_wall = (np.abs(mx) < 1e-12) | (np.abs(mx - 2.2) < 1e-12) | (np.abs(my) < 1e-12) | (np.abs(my - 0.41) < 1e-12)
    r_end = np.hypot(x[bnd] - 0.2, y[bnd] - 0.2)
    on_cyl = (np.abs(r_end - 0.05) < 1e-12).all(axis=1)
    assert (on_wall | on_cyl).all()
    assert on_cyl.sum() == m["n_cyl"]
    # polygonal hole: the cylinder edges are chords of the circle
    chord_area = 0.0
    for a, b in bnd[on_cyl]:
        chord_area += 0.5 * abs((x[a] - 0.2) * (y[b] - 0.2) - (x[b] - 0.2) * (y[a] - 0.2))
    assert abs(area.sum() - (2.2 * 0.41 - chord_area)) < 1e-10


def test_csr_graph(mesh):
    """Rows sorted, unique, contain the diagonal; the graph is the edge graph of the mesh."""
    _, m = mesh
    rp, ci, n = m["rp"], m["ci"], m["n"]
    ev = m["ev"].reshape(-1, 3)
    edges = set()
    for a, b in ((0, 1), (1, 2), (2, 0)):
        for i, j in zip(ev[:, a].tolist(), ev[:, b].tolist()):
            edges.add((i, j))
            edges.add((j, i))
    count = 0
    for i in range(n):
        row = ci[rp[i]:rp[i + 1]]
        assert (np.diff(row) > 0).all()
        assert i in row
        count += len(row) - 1
        for j in row.tolist():
            if j != i:
                assert (i, j) in edges
    assert count == len(edges)


def test_boundary_tags(mesh):
    """Tags as src/cfa/mesh.cuh classify_boundary_node: cylinder first, then walls (corners
    included), inlet, outlet; 1 inlet, 2 outlet, 3 walls, 4 cylinder, 0 interior."""
    _, m = mesh
    x, y, bc = m["x"], m["y"], m["bc"]
    cyl = np.abs(np.hypot(x - 0.2, y - 0.2) - 0.05) < 1e-9
    wall = (np.abs(y) < 1e-12) | (np.abs(y - 0.41) < 1e-12)
    inlet = (np.abs(x) < 1e-12) & ~wall
    outlet = (np.abs(x - 2.2) < 1e-12) & ~wall
    assert (bc[cyl] == 4).all()
    assert (bc[wall] == 3).all()
    assert (bc[inlet] == 1).all()
    assert (bc[outlet] == 2).all()
    assert (bc[~(cyl | wall | inlet | outlet)] == 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_device_nvb_uniform_matches_oracle(level):
    print(f"[case] uniform level {level}: device NVB forest vs the oracle", flush=True)
    d, o = cfa.uniform_mesh(level), cfa_oracle.uniform_mesh(level)
    assert (d["n"], d["ne"], d["nnz"]) == (o["n"], o["ne"], o["nnz"])
    if level <= 4:
        assert d["n"] == NODES[level] and d["nnz"] == NNZ[level]
    for key in ("ev", "bc", "rp", "ci", "cyl"):
        np.testing.assert_array_equal(d[key], o[key], err_msg=key)
    for key in ("x", "y"):
        np.testing.assert_array_equal(d[key], o[key], err_msg=key)
