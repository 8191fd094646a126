"""The solver-study dump format (--dump-systems N; src/cylinder_fem_amr.cu:331-352, 375-378,
426-438), on the CPU: the oracle writes the systems of one step of the MMS L1 run, the package's
reader parses them, and what the mathematics fixes is checked on the parsed systems:

* the pattern is the mesh's CSR (the product's own mesh of uniform level 1), x and y are the node
  coordinates, bc_tag the node tags;
* momentum Dirichlet rows (inlet, walls, cylinder; the outlet is do-nothing) are identity rows
  whose right-hand side equals the prescribed value, which the initial guess already holds;
* the appended solution satisfies the solver's stopping rule (relative residual below the
  tolerance, within a factor for the recursive-vs-true residual gap) and agrees with a direct
  sparse solve;
* the pressure matrix is symmetric; outlet rows are identity rows with zero right-hand side and
  zero initial guess.

The C reader's error paths (missing file, truncated file, wrong layout) need no GPU.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cfa
import cfa_oracle

BC_INLET, BC_OUTLET, BC_WALL, BC_CYLINDER = 1, 2, 3, 4


@pytest.fixture(scope="module")
def dumped(tmp_path_factory):
    d = tmp_path_factory.mktemp("dump")
    prefix = str(d / "l1")
    rows, _, _ = cfa_oracle.run_fixed(3, 1e-3, 1, 0, mms=True, prefix=prefix, dump_systems=3)
    assert len(rows) == 3
    return prefix


def csr(s):
    return sp.csr_matrix((s["val"], s["ci"], s["rp"]), shape=(s["n"], s["n"]))


def test_momentum_dump_layout_and_mesh(dumped):
    m = cfa.read_system(dumped + "_momentum.bin")
    mesh = cfa_oracle.uniform_mesh(1)
    assert m["n"] == mesh["n"] == 1912 and m["nnz"] == mesh["nnz"] == 12904
    np.testing.assert_array_equal(m["rp"], mesh["rp"])
    np.testing.assert_array_equal(m["ci"], mesh["ci"])
    np.testing.assert_array_equal(m["x"], mesh["x"])
    np.testing.assert_array_equal(m["y"], mesh["y"])
    np.testing.assert_array_equal(m["bc_tag"], mesh["bc"])
    assert "u" in m and "v" in m


def test_momentum_dump_dirichlet_rows_and_solution(dumped):
    m = cfa.read_system(dumped + "_momentum.bin")
    A = csr(m)
    dirichlet = np.where(np.isin(m["bc_tag"], (BC_INLET, BC_WALL, BC_CYLINDER)))[0]
    assert len(dirichlet) > 0
    for i in dirichlet:
        lo, hi = m["rp"][i], m["rp"][i + 1]
        row = dict(zip(m["ci"][lo:hi], m["val"][lo:hi]))
        assert row[i] == 1.0 and all(v == 0.0 for c, v in row.items() if c != i)
        assert m["b_u"][i] == m["u0"][i] == m["u"][i]
        assert m["b_v"][i] == m["v0"][i] == m["v"][i]
    for b, x0, x in ((m["b_u"], m["u0"], m["u"]), (m["b_v"], m["v0"], m["v"])):
        r0 = np.linalg.norm(b - A @ x0)
        r = np.linalg.norm(b - A @ x)
        assert r <= 10 * 1e-7 * r0, (r, r0)
        direct = spla.spsolve(A.tocsc(), b)
        assert np.max(np.abs(x - direct)) <= 1e-6 * np.max(np.abs(direct))


def test_pressure_dump_structure(dumped):
    p = cfa.read_system(dumped + "_pressure.bin")
    m = cfa.read_system(dumped + "_momentum.bin")
    K = csr(p)
    np.testing.assert_array_equal(p["rp"], m["rp"])
    assert abs(K - K.T).max() == 0.0
    outlet = m["bc_tag"] == BC_OUTLET
    assert outlet.sum() == 15
    for i in np.where(outlet)[0]:
        lo, hi = p["rp"][i], p["rp"][i + 1]
        row = dict(zip(p["ci"][lo:hi], p["val"][lo:hi]))
        assert row[i] == 1.0 and all(v == 0.0 for c, v in row.items() if c != i)
    assert np.all(p["b"][outlet] == 0.0) and np.all(p["x0"][outlet] == 0.0)
    # K_p restricted to the free rows is a stiffness matrix: positive diagonal, the system is SPD
    assert np.all(K.diagonal() > 0.0)
    x = spla.spsolve(K.tocsc(), p["b"])
    assert x @ (K @ x) > 0.0


def test_reader_rejects_bad_files(dumped, tmp_path):
    raw = open(dumped + "_momentum.bin", "rb").read()
    bad = tmp_path / "cut.bin"
    bad.write_bytes(raw[:-5])
    with pytest.raises(ValueError):
        cfa.read_system(bad)
    _, host = cfa.load()
    cfa._declare_dump_functions(host)
    n, nnz = ctypes.c_int(0), ctypes.c_int(0)
    assert host.cfa_dump_header((dumped + "_momentum.bin").encode(), ctypes.byref(n), ctypes.byref(nnz)) == 0
    assert (n.value, nnz.value) == (1912, 12904)
    CFA_E_IO = -4
    out = np.zeros(11)
    d = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    # a truncated file, a missing file and a pressure file read as momentum are rejected before
    # anything touches the device
    assert host.cfa_dumped_momentum_solve(str(bad).encode(), 1e-7, 150, 0, 0, None, None, d(out)) == CFA_E_IO
    assert host.cfa_dumped_momentum_solve(str(tmp_path / "none.bin").encode(), 1e-7, 150, 0, 0, None, None,
                                          d(out)) == CFA_E_IO
    assert host.cfa_dumped_momentum_solve((dumped + "_pressure.bin").encode(), 1e-7, 150, 0, 0, None, None,
                                          d(out)) == CFA_E_IO
    assert host.cfa_dumped_pressure_solve((dumped + "_momentum.bin").encode(), 1e-7, 100, 0, 0, 0, None,
                                          d(out)) == CFA_E_IO
    # a pattern that is not a CSR of an n x n matrix (a column index out of range)
    arr = bytearray(raw)
    n0 = 1912
    off_ci = 8 + 4 * (n0 + 1)
    arr[off_ci:off_ci + 4] = np.int32(n0 + 7).tobytes()
    bad2 = tmp_path / "badcol.bin"
    bad2.write_bytes(bytes(arr))
    assert host.cfa_dumped_momentum_solve(str(bad2).encode(), 1e-7, 150, 0, 0, None, None, d(out)) == CFA_E_IO
    # invalid arguments
    assert host.cfa_dumped_momentum_solve((dumped + "_momentum.bin").encode(), 0.0, 150, 0, 0, None, None,
                                          d(out)) == -1
