"""Discrete adjoint at the size of the gradient configuration: the frozen uniform level-4 mesh
(115,648 nodes), dt 1e-3, a 1000-step spin-up from the DFG initial condition, then J = mean C_D
over a 200-step window with checkpoints every 20 steps, every solve at 1e-12.

* dJ/dnu and dJ/dU_max from one backward pass against central differences of the forward window
  with relative step 1e-6: agreement to 1e-5 relative.
* dJ/deps for the 64 radial modes of the hole boundary from the same kind of pass, along a seeded
  random direction and for two single modes, against central differences of the forward window
  on deformed meshes with step 1e-6: agreement to 1e-5 relative (of the larger of the difference
  and a thousandth of the largest mode sensitivity).
The hierarchy is rebuilt on the same steps in every forward window, so the differences see the
same discrete map as the adjoint.
"""
import numpy as np
import pytest

import cfa

pytestmark = pytest.mark.gpu
THETA0 = np.array([1.0e-3, 1.5])
SPIN, W, EVERY = 1000, 200, 20
FIRST = SPIN + 1


@pytest.fixture(scope="module")
def adj4():
    print("[case] adjoint context level 4, tol 1e-12", flush=True)
    a = cfa.Adjoint(4, dt=1e-3, tol=1e-12)
    yield a
    a.close()


@pytest.fixture(scope="module")
def spun4(adj4):
    print(f"[case] level 4 spin-up, {SPIN} steps", flush=True)
    _, state, hist = adj4.forward(THETA0, 1, SPIN, adj4.initial_state())
    print(f"  C_D at step {SPIN}: {hist[-1, 0]:.10f}, C_L {hist[-1, 1]:.10f}", flush=True)
    assert np.all(np.isfinite(state))
    return state


def test_level4_theta_gradient_vs_central_differences(adj4, spun4):
    print(f"[case] level 4, {W}-step window: dJ/dtheta vs central differences", flush=True)
    J, g, _, st = adj4.gradient(THETA0, FIRST, W, spun4, EVERY)
    assert st["recomputed_steps"] == W
    fd = np.zeros(2)
    for k in range(2):
        h = 1e-6 * THETA0[k]
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (adj4.forward(tp, FIRST, W, spun4)[0] - adj4.forward(tm, FIRST, W, spun4)[0]) / (2 * h)
    print(f"  J {J:.12f}  adjoint {g}  central differences {fd}", flush=True)
    for k in range(2):
        assert abs(g[k] - fd[k]) <= 1e-5 * abs(fd[k]), (k, g[k], fd[k])


def test_level4_shape_gradient_vs_central_differences(adj4, spun4):
    print(f"[case] level 4, {W}-step window: dJ/deps (64 modes) vs central differences", flush=True)
    zero = np.zeros(64)
    adj4.set_shape(zero)
    try:
        _, _, _, st = adj4.gradient(THETA0, FIRST, W, spun4, EVERY, shape_gradient=True)
        g = st["dJ_dshape"]
        assert g.shape == (64,) and np.all(np.isfinite(g))
        rng = np.random.default_rng(11)
        h = 1e-6
        for d in (rng.standard_normal(64), np.eye(64)[0], np.eye(64)[16]):
            adj4.set_shape(h * d)
            jp = adj4.forward(THETA0, FIRST, W, spun4)[0]
            adj4.set_shape(-h * d)
            jm = adj4.forward(THETA0, FIRST, W, spun4)[0]
            fd = (jp - jm) / (2 * h)
            ad = float(g @ d)
            print(f"  <dJ/deps, d> adjoint {ad:.10e}  central difference {fd:.10e}", flush=True)
            assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (ad, fd)
    finally:
        adj4.set_shape(zero)
