"""The oracle's standalone solvers pinned by what the mathematics fixes (CPU only).

Systems: the oracle's --dump-systems files of step 1 on uniform level 1 (DFG inflow, dt 1e-3):
the non-symmetric momentum matrix (BDF2 mass, viscosity, SUPG convection, Dirichlet rows) and the
PPE matrix K_p with its outlet rows.

* Every solve lands on the sparse direct solution (scipy's SuperLU) within what its tolerance
  allows, and its true residual meets the stopping rule (relative to the initial residual).
* Jacobi PCG (src/cfa/linalg.cuh:172-256 with the inverse diagonal of :28-47) takes exactly the
  iterations of scipy's preconditioned CG with the same stopping rule ||r|| < tol ||b||: a wrong
  step length or direction update changes the count.
* Jacobi BiCGSTAB (src/cfa/linalg.cuh:261-347) is right-preconditioned like scipy's bicgstab; the
  counts agree to one (the reference counts the half-step exit on ||s||, scipy's callback does
  not), and the stopping rule decides: one iteration fewer reports no convergence.
* SA-AMG PCG and ILU(0)-BiCGSTAB2 land on the direct solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cfa
import cfa_oracle


@pytest.fixture(scope="module")
def systems(tmp_path_factory):
    prefix = str(tmp_path_factory.mktemp("oracle_solvers") / "u1")
    cfa_oracle.run_fixed(2, 1e-3, 1, 0, mms=False, prefix=prefix, dump_systems=1)
    m = cfa.read_system(prefix + "_momentum.bin")
    p = cfa.read_system(prefix + "_pressure.bin")
    A = sp.csr_matrix((m["val"], m["ci"], m["rp"]), shape=(m["n"], m["n"]))
    K = sp.csr_matrix((p["val"], p["ci"], p["rp"]), shape=(p["n"], p["n"]))
    return m, p, A, K


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("tol", [1e-7, 1e-10])
def test_bicgstab_jacobi(systems, tol):
    m, _, A, _ = systems
    b, x0 = m["b_u"], np.zeros(m["n"])
    x, st = cfa_oracle.bicgstab_jacobi(m["rp"], m["ci"], m["val"], b, x0, tol)
    assert st["converged"]
    assert np.linalg.norm(b - A @ x) <= 10 * tol * np.linalg.norm(b)
    assert rel(x, spla.spsolve(A.tocsc(), b)) <= 100 * tol
    count = [0]
    spla.bicgstab(A, b, rtol=tol, atol=0.0, M=sp.diags(1.0 / A.diagonal()), maxiter=200,
                  callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert abs(st["iterations"] - count[0]) <= 1, (st, count)
    # the stopping rule: with one iteration fewer the solve does not converge
    _, st_short = cfa_oracle.bicgstab_jacobi(m["rp"], m["ci"], m["val"], b, x0, tol, max_iter=st["iterations"] - 1)
    assert not st_short["converged"]


@pytest.mark.parametrize("tol", [1e-7, 1e-10])
def test_pcg_jacobi(systems, tol):
    _, p, _, K = systems
    b = p["b"]
    x, st = cfa_oracle.pcg_jacobi(p["rp"], p["ci"], p["val"], b, np.zeros(p["n"]), tol)
    assert st["converged"]
    assert np.linalg.norm(b - K @ x) <= 10 * tol * np.linalg.norm(b)
    assert rel(x, spla.spsolve(K.tocsc(), b)) <= 100 * tol
    count = [0]
    spla.cg(K, b, rtol=tol, atol=0.0, M=sp.diags(1.0 / K.diagonal()), maxiter=4000,
            callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert st["iterations"] == count[0], (st, count)


def test_pcg_amg_and_bicgstab2(systems):
    m, p, A, K = systems
    tol = 1e-10
    x, st = cfa_oracle.pcg_amg(p["rp"], p["ci"], p["val"], p["b"], np.zeros(p["n"]), tol)
    assert st["converged"] and st["amg_levels"] >= 2
    assert st["iterations"] < 60
    assert rel(x, spla.spsolve(K.tocsc(), p["b"])) <= 100 * tol
    u, v, st = cfa_oracle.bicgstab2_ilu(m["rp"], m["ci"], m["val"], m["b_u"], m["b_v"], m["u0"], m["v0"], tol)
    assert st["converged_u"] and st["converged_v"] and st["colours"] >= 3
    assert rel(u, spla.spsolve(A.tocsc(), m["b_u"])) <= 100 * tol
    assert np.linalg.norm(m["b_v"] - A @ v) <= 1e-11 + 10 * tol * np.linalg.norm(m["b_v"] - A @ m["v0"])
