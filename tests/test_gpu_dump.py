"""--dump-systems on the GPU: the program's dump of one step against the oracle's dump of the same
step, and the dumped systems solved again through the C ABI (cfa_dumped_momentum_solve,
cfa_dumped_pressure_solve).

* Step 1 of the flow past the cylinder on uniform level 1 (DFG inflow, dt 1e-3): the momentum
  system is assembled from the initial condition only, so the matrix, both right-hand sides and
  the initial guesses (with their Dirichlet values) are bit-exact against the oracle, as are the
  pattern, coordinates and tags.
* The appended solution and the pressure system depend on a Krylov solve (the order of the
  reductions differs from the oracle's), so they follow the oracle to 1e-10 relative.
* Solving the dumped momentum system again from its dumped initial guess with the same solver
  and tolerance reproduces the dumped solution (1e-13 relative); the FP32 factors converge to
  the same solution within the tolerance.
* The dumped pressure system (K_p alone) solved with SA-AMG PCG from zero meets its stopping rule
  in a residual computed on the host, in FP64 and with the FP32 cycle.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cfa
import cfa_oracle

pytestmark = pytest.mark.gpu


def rel_dev(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    d = tmp_path_factory.mktemp("gpudump")
    gpu, ora = str(d / "gpu"), str(d / "ora")
    print("[case] uniform level 1, dump of step 1 (GPU run and oracle)", flush=True)
    assert cfa.run(["--uniform-level", "1", "--dt", "0.001", "--steps", "2", "--min-level", "0",
                    "--quiet", "--no-csv", "--dump-systems", "1", "--prefix", gpu]) == 0
    cfa_oracle.run_fixed(2, 1e-3, 1, 0, mms=False, prefix=ora, dump_systems=1)
    return gpu, ora


def test_momentum_dump_matches_oracle(dumps):
    gpu, ora = (cfa.read_system(p + "_momentum.bin") for p in dumps)
    for key in ("rp", "ci", "x", "y", "bc_tag", "val", "b_u", "b_v", "u0", "v0"):
        np.testing.assert_array_equal(gpu[key], ora[key], err_msg=key)
    for key in ("u", "v"):
        assert rel_dev(gpu[key], ora[key]) <= 1e-10, key


def test_pressure_dump_matches_oracle(dumps):
    gpu, ora = (cfa.read_system(p + "_pressure.bin") for p in dumps)
    np.testing.assert_array_equal(gpu["rp"], ora["rp"])
    np.testing.assert_array_equal(gpu["ci"], ora["ci"])
    for key in ("val", "b", "x0"):
        assert rel_dev(gpu[key], ora[key]) <= 1e-10, key


def test_dumped_momentum_solved_again(dumps):
    path = dumps[0] + "_momentum.bin"
    print("[case] dumped momentum system, BiCGSTAB2 + ILU(0), FP64 and FP32 factors", flush=True)
    st, u, v = cfa.solve_dumped_momentum(path, tol=1e-7)
    assert st["converged_u"] == 1.0 and st["converged_v"] == 1.0
    assert 0.0 <= st["rel_dev_from_dumped"] <= 1e-13, st
    m = cfa.read_system(path)
    assert rel_dev(u, m["u"]) <= 1e-13 and rel_dev(v, m["v"]) <= 1e-13
    st32, u32, v32 = cfa.solve_dumped_momentum(path, tol=1e-7, precision="mixed")
    assert st32["converged_u"] == 1.0 and st32["converged_v"] == 1.0
    A = sp.csr_matrix((m["val"], m["ci"], m["rp"]), shape=(m["n"], m["n"]))
    for b, x0, x in ((m["b_u"], m["u0"], u32), (m["b_v"], m["v0"], v32)):
        assert np.linalg.norm(b - A @ x) <= 10 * 1e-7 * np.linalg.norm(b - A @ x0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dumped_pressure_solved_from_zero(dumps, precision):
    path = dumps[0] + "_pressure.bin"
    print(f"[case] dumped pressure system, SA-AMG PCG ({precision}) from 0", flush=True)
    st, x = cfa.solve_dumped_pressure(path, tol=1e-8, from_zero=True, precision=precision)
    assert st["converged"] == 1.0 and st["amg_levels"] >= 2, st
    p = cfa.read_system(path)
    K = sp.csr_matrix((p["val"], p["ci"], p["rp"]), shape=(p["n"], p["n"]))
    assert np.linalg.norm(p["b"] - K @ x) <= 10 * 1e-8 * np.linalg.norm(p["b"])
