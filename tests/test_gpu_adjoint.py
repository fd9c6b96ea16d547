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
    # the grid force cotangent is accumulated with FP64 atomics: bits may
    # differ from call to call in the last places
    assert torch.allclose(th.grad, torch.as_tensor(2.0 * grad), rtol=1e-9, atol=0)
