"""Adjoint gradient of the dipole misfit (hv_dipole_loss_grad) against
central finite differences (relative step 1e-4) in FP64, on the A=40 nucleus
with 4,000 test particles.  The SOR is run to a tight tolerance here so that
the forward map is the exact-solve map the adjoint differentiates.  The loss
is only piecewise smooth (CIC weights are piecewise linear in the positions),
so finite differences carry an error of their own that falls with the number
of test particles: with 1,000 test particles a relative step of 1e-4 lands
within 2e-3 of the adjoint, with 4,000 within 2.5e-4."""
import numpy as np
import pytest

import hvamd
from hvamd import configs
from hvamd.adjoint import DipoleFit, DipoleLoss, central_difference, dipole_series

pytestmark = pytest.mark.gpu


def _setup(nsteps=40, ck=16):
    w = configs.c0()
    prm = w.params(nsteps)
    prm.sor_tol = 1e-11
    prm.sor_max_it = 4000
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    target = dipole_series(w, theta * 1.03, nsteps, prm)
    return w, prm, theta, target, nsteps, ck


def test_adjoint_matches_central_differences():
    print("adjoint vs FD", flush=True)
    w, prm, theta, target, nsteps, ck = _setup()
    fit = DipoleFit(w, target, nsteps, ck, prm)
    loss, grad, d = fit.loss_grad(theta)
    fd = central_difference(fit, theta, 1e-4)
    loss2, _, d2 = fit.loss_grad(theta, grad=False)
    fit.close()
    print("loss", loss, "adjoint", grad, "fd", fd)
    assert loss == loss2 and np.array_equal(d, d2)  # repeatable: the call restores the initial particles
    assert loss > 0
    for k in range(4):
        assert abs(grad[k] - fd[k]) <= 1e-3 * abs(fd[k]) + 1e-12 * abs(loss), (k, grad[k], fd[k])


def test_autograd_function_backward_is_the_adjoint():
    torch = pytest.importorskip("torch")
    print("autograd", flush=True)
    w, prm, theta, target, nsteps, ck = _setup(24, 10)
    fit = DipoleFit(w, target, nsteps, ck, prm)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    loss = DipoleLoss.apply(th, fit)
    (2.0 * loss).backward()
    _, grad, _ = fit.loss_grad(theta)
    fit.close()
    # the gradient is deterministic (the grid-force cotangent is an exact
    # fixed-point sum), so the backward equals a second adjoint call bit for bit
    assert torch.equal(th.grad, torch.as_tensor(2.0 * grad))


def test_adjoint_gradient_is_deterministic():
    """Two calls give the same gradient bits (the gathered cotangents are summed
    as exact integers, whatever order the atomics land in)."""
    print("adjoint determinism", flush=True)
    w, prm, theta, target, nsteps, ck = _setup(30, 12)
    fit = DipoleFit(w, target, nsteps, ck, prm)
    try:
        g1 = fit.loss_grad(theta)[1]
        g2 = fit.loss_grad(theta)[1]
    finally:
        fit.close()
    assert np.array_equal(g1, g2), (g1, g2)


def test_adjoint_c4_workload_200_steps():
    """The differentiation workload itself (A=120, 120,000 test particles,
    64^3, dx 0.4 fm, dt 0.5 fm/c; target from theta x 1.03; SOR to 1e-11) over
    200 steps with checkpoints every 100: the adjoint gradient within 1e-3 of
    central differences at a relative step of 1e-5 in every component.  The
    loss is only piecewise smooth (piecewise-linear CIC weights) and strongly
    curved in gamma: at a step of 1e-4 the finite difference of gamma itself
    is off by ~1e-2; at 1e-5 it agrees with the adjoint to 7e-4, the others to
    1e-4.  Both the forward run and the adjoint are deterministic, so the
    figures repeat exactly.  Past a few hundred steps the chaos of the
    test-particle dynamics separates the exact derivative of the discrete map
    from any finite secant."""
    print("adjoint c4 200 steps vs FD h=1e-5", flush=True)
    w = configs.c4()
    nsteps = 200
    prm = w.params(nsteps)
    prm.sor_tol = 1e-11
    prm.sor_max_it = 4000
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    target = dipole_series(w, theta * 1.03, nsteps, prm)
    fit = DipoleFit(w, target, nsteps, 100, prm)
    try:
        loss, grad, _ = fit.loss_grad(theta)
        fd = central_difference(fit, theta, 1e-5)
    finally:
        fit.close()
    rel = np.abs(grad - fd) / np.abs(fd)
    print("loss", loss, "adjoint", grad, "fd", fd, "relative", rel)
    assert loss > 0
    assert rel.max() < 1e-3, rel


def test_adjoint_c4_full_horizon_windowed():
    """The differentiation workload over its whole 1,500 steps as a windowed
    misfit (hvamd.adjoint.WindowedDipoleFit: 15 windows of 100 steps, each
    started from the target trajectory's state, multiple shooting; SOR to
    1e-11): the adjoint gradient's derivative along theta itself against the
    central difference of the same summed misfit at the workload's relative
    step 1e-4, within 1e-3 (per component, measured 7e-5 at h = 1e-5 and 3e-5,
    3e-4 at 1e-4)."""
    from hvamd.adjoint import WindowedDipoleFit, target_run
    print("adjoint c4 1500 steps windowed vs directional FD h=1e-4", flush=True)
    w = configs.c4()
    nsteps, window = w.steps, 100
    prm = w.params(nsteps)
    prm.sor_tol = 1e-11
    prm.sor_max_it = 4000
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    target, states = target_run(w, theta * 1.03, nsteps, window, prm)
    fit = WindowedDipoleFit(w, target, states, window, 100, prm)
    try:
        loss, grad, _ = fit.loss_grad(theta)
        print("loss", loss, "adjoint", grad, flush=True)
        h = 1e-4
        lp = fit.loss_grad(theta * (1.0 + h), grad=False)[0]
        lm = fit.loss_grad(theta * (1.0 - h), grad=False)[0]
    finally:
        fit.close()
    fd = (lp - lm) / (2.0 * h)
    ad = float(np.dot(grad, theta))
    print("directional: adjoint", ad, "fd", fd, "relative", (ad - fd) / abs(fd))
    assert len(states) == 15 and loss > 0
    assert abs(ad - fd) < 1e-3 * abs(fd)
