"""Discrete adjoint of the projection step on the GPU (C ABI cfa_step_fwd / cfa_window_forward /
cfa_window_adjoint), on the frozen uniform level-1 mesh (1912 nodes), dt 1e-3.

* The adjoint context's forward step is the program's step: from the DFG initial condition, 60
  steps at the program's tolerances (1e-7) give the program's C_D, C_L history bit for bit
  (same kernels, same order, AMG rebuilt at steps 1 and 51).
* dJ/dtheta for theta = (nu, U_max), J the mean C_D over a 12-step window after a 30-step
  spin-up (checkpoints every 5 steps), against central finite differences of the forward window
  with relative step 1e-6, both at solver tolerance 1e-12: agreement to 1e-5 relative.
* dJ/dstate0 in a seeded random direction against a central difference (Taylor check).
* dJ/deps for the 64 radial modes of the hole boundary (shape sensitivities through every
  element kernel's coordinates) against central differences of the forward window on deformed
  meshes, along a seeded random direction and for single modes, relative step 1e-5: 1e-5.
* The gradient does not depend on the checkpoint spacing.
* A torch.autograd.Function (this harness only) whose forward is cfa_window_forward and whose
  backward is cfa_window_adjoint gives the same gradient through theta.grad.
* On the level-2 mesh (7408 nodes), J = mean C_D over the first 12 steps from the initial
  condition (theta acting on every step): the GPU window equals the oracle's run (the serial
  FP64 CPU restatement) to 1e-9, and the GPU adjoint gradient matches central differences of the
  oracle's J to 1e-5 relative.
"""
import numpy as np
import pytest
import torch

import cfa
import cfa_oracle

pytestmark = pytest.mark.gpu
THETA0 = np.array([1.0e-3, 1.5])
SPIN, FIRST, W, EVERY = 30, 31, 12, 5


@pytest.fixture(scope="module")
def adj():
    print("[case] adjoint context level 1, tol 1e-12", flush=True)
    a = cfa.Adjoint(1, dt=1e-3, tol=1e-12)
    yield a
    a.close()


@pytest.fixture(scope="module")
def spun(adj):
    _, state, _ = adj.forward(THETA0, 1, SPIN, adj.initial_state())
    return state


class WindowForce(torch.autograd.Function):
    """J(theta) over a window: forward through cfa_window_forward, backward through the
    discrete adjoint cfa_window_adjoint."""

    @staticmethod
    def forward(fctx, theta, adj, state0, first_step, steps, every):
        J, _, _ = adj.forward(theta.detach().numpy(), first_step, steps, state0)
        fctx.adj, fctx.state0, fctx.args = adj, state0, (first_step, steps, every)
        fctx.save_for_backward(theta)
        return theta.new_tensor(J)

    @staticmethod
    def backward(fctx, grad_J):
        (theta,) = fctx.saved_tensors
        first_step, steps, every = fctx.args
        _, g, _, _ = fctx.adj.gradient(theta.detach().numpy(), first_step, steps, fctx.state0, every)
        return grad_J * torch.from_numpy(g), None, None, None, None, None


def test_window_forward_is_the_program_step():
    print("[case] adjoint forward vs program, level 1, 60 steps, tol 1e-7", flush=True)
    a = cfa.Adjoint(1, dt=1e-3, tol=1e-7)
    _, _, hist = a.forward(THETA0, 1, 60, a.initial_state())
    a.close()
    assert cfa.run(["--uniform-level", "1", "--min-level", "0", "--dt", "0.001", "--steps", "60", "--quiet",
                    "--no-csv"]) == 0
    H = {c: i for i, c in enumerate(cfa.HISTORY_COLUMNS)}
    prog = cfa.last_history()
    np.testing.assert_array_equal(hist[:, 0], prog[:, H["C_D"]])
    np.testing.assert_array_equal(hist[:, 1], prog[:, H["C_L"]])


def test_step_matches_window(adj, spun):
    s1, f1 = adj.step(THETA0, FIRST, spun)
    J, s_w, hist = adj.forward(THETA0, FIRST, 1, spun)
    np.testing.assert_array_equal(s1, s_w)
    assert f1[0] == hist[0, 0] == J


def test_theta_gradient_vs_central_differences(adj, spun):
    print("[case] dJ/dtheta vs central differences", flush=True)
    J, g, _, st = adj.gradient(THETA0, FIRST, W, spun, EVERY)
    J0, _, _ = adj.forward(THETA0, FIRST, W, spun)
    assert abs(J - J0) <= 1e-12 * abs(J0)
    assert st["recomputed_steps"] == W
    fd = np.zeros(2)
    for k in range(2):
        h = 1e-6 * THETA0[k]
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (adj.forward(tp, FIRST, W, spun)[0] - adj.forward(tm, FIRST, W, spun)[0]) / (2 * h)
    print(f"  J {J:.12f}  adjoint {g}  central differences {fd}", flush=True)
    for k in range(2):
        assert abs(g[k] - fd[k]) <= 1e-5 * abs(fd[k]), (k, g[k], fd[k])


def test_state_gradient_taylor(adj, spun):
    print("[case] dJ/dstate0 along a seeded direction", flush=True)
    J, _, gs, _ = adj.gradient(THETA0, FIRST, W, spun, EVERY, state_gradient=True)
    n = adj.n
    rng = np.random.default_rng(7)
    d = np.zeros(6 * n)
    d[:5 * n] = rng.standard_normal(5 * n)
    eps = 1e-6 * np.max(np.abs(spun[:5 * n]))
    fd = (adj.forward(THETA0, FIRST, W, spun + eps * d)[0] - adj.forward(THETA0, FIRST, W, spun - eps * d)[0]) / (2 * eps)
    ad = float(gs @ d)
    print(f"  <dJ/dS0, d> adjoint {ad:.10e}  central difference {fd:.10e}", flush=True)
    assert np.all(gs[5 * n:] == 0.0)
    assert abs(ad - fd) <= 1e-5 * abs(fd)


def test_checkpoint_spacing_does_not_change_the_gradient(adj, spun):
    ref = adj.gradient(THETA0, FIRST, W, spun, 0)[1]
    for every in (1, 4, 7):
        g = adj.gradient(THETA0, FIRST, W, spun, every)[1]
        assert np.all(np.abs(g - ref) <= 1e-8 * np.abs(ref)), (every, g, ref)


def test_torch_autograd_function(adj, spun):
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    J = WindowForce.apply(theta, adj, spun, FIRST, W, EVERY)
    J.backward()
    _, g, _, _ = adj.gradient(THETA0, FIRST, W, spun, EVERY)
    np.testing.assert_array_equal(theta.grad.numpy(), g)
    assert abs(J.item() - adj.forward(THETA0, FIRST, W, spun)[0]) == 0.0


def test_level2_gradient_vs_oracle_central_differences():
    print("[case] level 2, 12 steps from the initial condition: GPU adjoint vs oracle differences", flush=True)
    steps = 12
    a = cfa.Adjoint(2, dt=1e-3, tol=1e-12)
    J, g, _, _ = a.gradient(THETA0, 1, steps, a.initial_state(), 4)
    a.close()

    def oracle_J(theta):
        return float(np.mean(cfa_oracle.run_params(steps, 1e-3, 2, theta[0], theta[1], 1e-12)[:, 2]))

    J_o = oracle_J(THETA0)
    assert abs(J - J_o) <= 1e-9 * abs(J_o), (J, J_o)
    fd = np.zeros(2)
    for k in range(2):
        h = 1e-6 * THETA0[k]
        tp, tm = THETA0.copy(), THETA0.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (oracle_J(tp) - oracle_J(tm)) / (2 * h)
    print(f"  J {J:.12f}  adjoint {g}  oracle central differences {fd}", flush=True)
    for k in range(2):
        assert abs(g[k] - fd[k]) <= 1e-5 * abs(fd[k]), (k, g[k], fd[k])


def test_shape_gradient_vs_central_differences(adj, spun):
    print("[case] dJ/deps (64 radial boundary modes) vs central differences", flush=True)
    zero = np.zeros(64)
    adj.set_shape(zero)
    try:
        _, _, _, st = adj.gradient(THETA0, FIRST, W, spun, EVERY, shape_gradient=True)
        g = st["dJ_dshape"]
        assert g.shape == (64,) and np.all(np.isfinite(g))
        rng = np.random.default_rng(11)
        directions = [rng.standard_normal(64), np.eye(64)[0], np.eye(64)[16]]
        h = 1e-5
        for d in directions:
            adj.set_shape(h * d)
            jp = adj.forward(THETA0, FIRST, W, spun)[0]
            adj.set_shape(-h * d)
            jm = adj.forward(THETA0, FIRST, W, spun)[0]
            fd = (jp - jm) / (2 * h)
            ad = float(g @ d)
            print(f"  <dJ/deps, d> adjoint {ad:.10e}  central difference {fd:.10e}", flush=True)
            assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (ad, fd)
    finally:
        adj.set_shape(zero)
