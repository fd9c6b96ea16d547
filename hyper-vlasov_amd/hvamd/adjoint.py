"""Gradient of a dipole-response misfit with respect to the Skyrme
coefficients (hv_dipole_loss_grad: the discrete adjoint of the FP64 loop),
wrapped as a torch.autograd.Function for fitting scripts.

    fit = DipoleFit(workload, target, nsteps=1500, checkpoint_every=100)
    theta = torch.tensor([A, B, C, gamma], dtype=torch.float64, requires_grad=True)
    loss = DipoleLoss.apply(theta, fit)
    loss.backward()            # theta.grad = dloss/d(A, B, C, gamma)

The loss is sum_s (D(t_s) - D_target(s))^2 over the steps of one run from the
workload's seeded initial state, D = N Z / A (<z>_p - <z>_n)
(scripts/programs.py:99).  central_difference() gives the finite-difference
gradient the adjoint is checked against."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import HV_PREC_F64
from .sim import Simulation, Skyrme

PARAM_NAMES = ("A", "B", "C", "gamma")


class DipoleFit:
    """Loss and gradient over a fixed initial state (FP64 context)."""

    def __init__(self, workload, target, nsteps: int, checkpoint_every: int = 100, params=None):
        self.w = workload
        self.nsteps = int(nsteps)
        self.ck = int(checkpoint_every)
        self.params = params if params is not None else workload.params(self.nsteps)
        self.target = np.ascontiguousarray(target, dtype=np.float64)
        if self.target.shape != (self.nsteps,):
            raise ValueError(f"target must have {self.nsteps} values")
        self.sim = Simulation(workload.world(), workload.n_protons, workload.n_neutrons, HV_PREC_F64)
        self.sim.upload_particles(*workload.initial_state())
        self.evaluations = 0

    def close(self):
        self.sim.close()

    def loss_grad(self, theta, grad: bool = True):
        """(loss, gradient[4] or None, D series) at theta = (A, B, C, gamma)."""
        sk = Skyrme(*[float(v) for v in theta]).c()
        pr = self.params.c()
        loss = ctypes.c_double()
        d = np.zeros(self.nsteps)
        g = np.zeros(4) if grad else None
        dp = _lib._DP
        rc = self.sim.lib.hv_dipole_loss_grad(self.sim.ctx, ctypes.byref(sk), ctypes.byref(pr), self.nsteps,
                                              self.target.ctypes.data_as(dp), self.ck, ctypes.byref(loss),
                                              d.ctypes.data_as(dp), g.ctypes.data_as(dp) if grad else None)
        self.sim._rc("hv_dipole_loss_grad", rc)
        self.evaluations += 1
        return loss.value, g, d


def dipole_series(workload, theta, nsteps: int, params=None) -> np.ndarray:
    """D(t_s) of a run at theta (e.g. the target of a fit)."""
    fit = DipoleFit(workload, np.zeros(nsteps), nsteps, max(nsteps, 1), params)
    try:
        return fit.loss_grad(theta, grad=False)[2]
    finally:
        fit.close()


def central_difference(fit: DipoleFit, theta, h: float = 1e-4) -> np.ndarray:
    """Central finite differences with relative steps h |theta_k|."""
    theta = np.asarray(theta, dtype=np.float64)
    out = np.zeros(4)
    for k in range(4):
        step = h * abs(theta[k])
        tp, tm = theta.copy(), theta.copy()
        tp[k] += step
        tm[k] -= step
        out[k] = (fit.loss_grad(tp, grad=False)[0] - fit.loss_grad(tm, grad=False)[0]) / (2.0 * step)
    return out


try:
    import torch

    class DipoleLoss(torch.autograd.Function):
        """loss(theta) with the adjoint gradient as its backward."""

        @staticmethod
        def forward(ctx, theta, fit):
            loss, grad, _ = fit.loss_grad(theta.detach().cpu().numpy())
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype))
            return torch.tensor(loss, dtype=theta.dtype)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g * grad, None
except ImportError:  # torch is optional for the C-ABI users
    DipoleLoss = None
