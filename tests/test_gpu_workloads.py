"""Isolated solver workloads (the path's PCG + SA-AMG and ILU-BiCGSTAB2 on frozen uniform meshes)
at a small level: residual reached, iteration counts sane, FP32 and FP64 preconditioners give the
same solution to the solver tolerance (the pressure solves with the FP32 cycle meet the same rule
within two more iterations per solve)."""
import pytest

import cfa

pytestmark = pytest.mark.gpu


def test_pressure_workload_level3():
    print("[case] pressure workload, level 3, 2 right-hand sides", flush=True)
    w = cfa.workload_pressure(3, nrhs=2, tol=1e-8)
    assert w["n"] == 29152 and w["solves"] == 2
    assert w["max_rel_residual"] < 1e-8
    assert 5 <= w["iterations"] / w["solves"] <= 200
    assert w["amg_levels"] >= 3
    # the FP32 cycle inside the FP64 PCG: the same stopping rule met, at most a few more iterations
    assert w["fp32_max_rel_residual"] < 1e-8
    assert w["fp32_iterations"] <= w["iterations"] + 2 * w["solves"]


def test_momentum_workload_level3():
    print("[case] momentum workload, level 3, FP64 and FP32 factors", flush=True)
    w = cfa.workload_momentum(3, tol=1e-8)
    assert w["n"] == 29152 and w["colours"] >= 3
    for k in ("fp64_it_u", "fp64_it_v", "fp32_it_u", "fp32_it_v"):
        assert 1 <= w[k] < 150, k
    assert w["fp32_it_u"] <= w["fp64_it_u"] + 2 and w["fp32_it_v"] <= w["fp64_it_v"] + 2
    assert w["rel_diff_fp32_fp64"] < 1e-6
