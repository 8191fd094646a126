"""The solver workloads and the mixed-precision run at their own sizes, on the GPU.

* Frozen uniform level 6 (1,838,848 nodes, 3,670,016 triangles, nnz 12,856,576), pressure: the PPE
  Laplacian (stab 0, outlet rows Dirichlet) with one SA-AMG set-up and PCG from 0 to 1e-8 for the
  seeded zero-mean right-hand sides mt19937_64(1000 + k), N(0,1). Every solution meets the
  stopping rule in a residual recomputed on the host from the downloaded matrix, and the first one
  is compared with the oracle's FP64 SA-AMG PCG on the same system (iterations within one,
  solution within 1e-8 relative).
* The same mesh, momentum: the BDF2 momentum matrix (dt 1e-3, nu 1e-3, SUPG, skew 0.5) for the
  seeded divergence-free field of amplitude 1.5, ILU(0)-BiCGSTAB2 to 1e-8 with FP64 and with FP32
  factors, against the oracle's FP64 ILU(0)-BiCGSTAB2 on the same system: iterations within three
  of the oracle's at 1e-8 (the residual of BiCGSTAB is not monotone, so over about 45 iterations
  the reduction order can move the stopping step by two), and the FP64- and FP32-factor solutions
  within 5e-8 relative of the oracle's solution at 1e-12 (two iterates that met the 1e-8 rule
  differ by about the tolerance: the oracle's own 1e-8 iterate is held to the same bound); the
  FP32-factor solution meets the stopping rule in a residual recomputed on the host.
* The adaptive flow with refinement capped at level 6 (levels 2-6, element budget 600k, remesh
  every 20 steps overlapped by 10, dt 5e-4, tolerances 1e-8), 300 steps in FP64 and with the FP32
  preconditioners: the meshes must stay identical step by step, and the force coefficients must
  agree within 1e-8 from step 100 on (the impulsive start, C_D near 276 at step 1, is reported).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cfa
import cfa_oracle

pytestmark = pytest.mark.gpu
H = {c: i for i, c in enumerate(cfa.HISTORY_COLUMNS)}
U6 = {"n": 1838848, "nnz": 12856576, "ne": 3670016}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def u6_pressure():
    print("[case] uniform level 6 pressure workload, 4 right-hand sides", flush=True)
    cfa.workload_keep_systems(True)
    try:
        w = cfa.workload_pressure(6, nrhs=4, tol=1e-8)
        s = cfa.workload_system("pressure")
    finally:
        cfa.workload_keep_systems(False)
    print(f"  set-up {w['amg_setup_ms']:.1f} ms, {w['amg_levels']:.0f} levels, {w['ms_per_solve']:.1f} ms per solve, "
          f"iterations {s['iters'].tolist()}", flush=True)
    return w, s


@pytest.fixture(scope="module")
def u6_momentum():
    print("[case] uniform level 6 momentum workload, FP64 and FP32 factors", flush=True)
    cfa.workload_keep_systems(True)
    try:
        w = cfa.workload_momentum(6, tol=1e-8)
        s = cfa.workload_system("momentum")
    finally:
        cfa.workload_keep_systems(False)
    print(f"  colours {w['colours']:.0f}, FP64 {w['fp64_solve_ms']:.1f} ms, FP32 {w['fp32_solve_ms']:.1f} ms, "
          f"iterations {s['iters'].tolist()}", flush=True)
    return w, s


def test_u6_pressure_residuals_on_host(u6_pressure):
    w, s = u6_pressure
    assert (w["n"], w["nnz"], w["ne"]) == (U6["n"], U6["nnz"], U6["ne"])
    assert w["max_rel_residual"] < 1e-8 and w["amg_levels"] >= 5
    K = sp.csr_matrix((s["val"], s["ci"], s["rp"]), shape=(s["n"], s["n"]))
    outlet = s["tag"] == 2
    for k in range(s["b"].shape[0]):
        b, x = s["b"][k], s["x"][k]
        assert abs(b[~outlet].sum()) < 1e-8 * np.abs(b).sum() and not b[outlet].any()
        r = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
        print(f"  right-hand side {k}: {s['iters'][k]} iterations, host residual {r:.3e}", flush=True)
        # the recursive residual met 1e-8; the recomputed one may carry the rounding of the updates
        assert r <= 2e-8, (k, r)


def test_u6_pressure_vs_oracle(u6_pressure):
    _, s = u6_pressure
    x, st = cfa_oracle.pcg_amg(s["rp"], s["ci"], s["val"], s["b"][0], np.zeros(s["n"]), tol=1e-8)
    d = rel(s["x"][0], x)
    print(f"  oracle: {st['iterations']} iterations ({st['amg_levels']} levels); GPU {s['iters'][0]}; "
          f"max rel deviation {d:.3e}", flush=True)
    assert st["converged"]
    assert abs(int(s["iters"][0]) - st["iterations"]) <= 1
    assert d <= 1e-8


def test_u6_momentum_vs_oracle(u6_momentum):
    w, s = u6_momentum
    assert (w["n"], w["nnz"]) == (U6["n"], U6["nnz"])
    n = s["n"]
    u, v, st = cfa_oracle.bicgstab2_ilu(s["rp"], s["ci"], s["val"], s["b"][0], s["b"][1], s["x0"][0], s["x0"][1],
                                        tol=1e-8)
    assert st["converged_u"] and st["converged_v"] and st["colours"] == int(w["colours"])
    it64 = s["iters"][:2].tolist()
    print(f"  oracle iterations ({st['it_u']}, {st['it_v']}), GPU FP64 {it64}, FP32 {s['iters'][2:].tolist()}",
          flush=True)
    # BiCGSTAB's residual is not monotone: over 45 iterations the reduction order moves the
    # first step below the tolerance by up to a few iterations
    assert abs(it64[0] - st["it_u"]) <= 3 and abs(it64[1] - st["it_v"]) <= 3
    # an iterate that met the 1e-8 residual rule is within about cond(A) * 1e-8 of the solution:
    # the GPU's iterates are held to three times the oracle's own 1e-8 iterate's distance from the
    # oracle's solution at 1e-12, and their residuals are recomputed on the host
    uc, vc, stc = cfa_oracle.bicgstab2_ilu(s["rp"], s["ci"], s["val"], s["b"][0], s["b"][1], s["x0"][0],
                                           s["x0"][1], tol=1e-12)
    assert stc["converged_u"] and stc["converged_v"]
    A = sp.csr_matrix((s["val"], s["ci"], s["rp"]), shape=(n, n))
    for c, exact, o8 in ((0, uc, u), (1, vc, v)):
        e64, e32, eo = rel(s["x"][c], exact), rel(s["x"][2 + c], exact), rel(o8, exact)
        r0 = np.linalg.norm(s["b"][c] - A @ s["x0"][c])
        r64 = np.linalg.norm(s["b"][c] - A @ s["x"][c]) / r0
        r32 = np.linalg.norm(s["b"][c] - A @ s["x"][2 + c]) / r0
        print(f"  component {c}: error vs the oracle at 1e-12: GPU FP64 factors {e64:.3e}, FP32 factors {e32:.3e}, "
              f"oracle at 1e-8 {eo:.3e}; host residuals FP64 {r64:.3e}, FP32 {r32:.3e}", flush=True)
        assert eo <= 1e-6
        assert e64 <= 3 * eo and e32 <= 3 * eo
        assert r64 <= 2e-8 and r32 <= 2e-8


@pytest.mark.parametrize("steps", [300])
def test_adaptive_level6_mixed_vs_fp64(steps):
    flags = ["--steps", str(steps), "--dt", "0.0005", "--amr-interval", "20", "--amr-overlap", "10",
             "--amr-tol-vort", "0.05", "--amr-tol-interp", "0.005", "--max-level", "6", "--max-elems", "600000",
             "--min-level", "2", "--mom-tol", "1e-8", "--ppe-tol", "1e-8", "--quiet", "--no-csv"]
    hist = {}
    for prec in ("fp64", "mixed"):
        print(f"[case] adaptive level 6, {steps} steps, precision {prec}", flush=True)
        assert cfa.run(flags + ["--precision", prec]) == 0
        hist[prec] = cfa.last_history()
        st = cfa.last_stats()
        print(f"  wall {st['wall_s']:.2f} s, final mesh {int(hist[prec][-1, H['nodes']])} nodes / "
              f"{int(hist[prec][-1, H['elements']])} elements", flush=True)
    a, b = hist["fp64"], hist["mixed"]
    np.testing.assert_array_equal(a[:, H["nodes"]], b[:, H["nodes"]])
    np.testing.assert_array_equal(a[:, H["elements"]], b[:, H["elements"]])
    assert a[-1, H["elements"]] > a[0, H["elements"]]
    for c in ("C_D", "C_L"):
        d = np.abs(a[:, H[c]] - b[:, H[c]])
        print(f"  max |d{c}| over all steps {d.max():.3e}, from step 100 {d[99:].max():.3e}", flush=True)
        assert d[99:].max() <= 1e-8, c
        assert d.max() <= 1e-7, c
