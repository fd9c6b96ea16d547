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
import json

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
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype, device=theta.device))
            return torch.tensor(loss, dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g * grad, None
except ImportError:  # torch is optional for the C-ABI users
    DipoleLoss = None


def study(name: str = "c4", horizons=(None,), hs=(1e-4,), checkpoint_every: int = 100,
          sor_tol: float | None = None, sor_max_it: int | None = None) -> dict:
    """The differentiation workload: for each horizon (steps of the run), the
    target from the run at theta * 1.03, the adjoint gradient at theta, and
    central differences at each relative step h.  The spread of the finite
    differences over h separates their own error (the loss is only piecewise
    smooth: CIC weights are piecewise linear in the positions) from the
    adjoint's."""
    import time

    from . import configs
    w = configs.WORKLOADS[name]()
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    out = {"workload": w.name, "test_particles": w.total, "checkpoint_every": checkpoint_every, "horizons": []}
    for nsteps in horizons:
        nsteps = w.steps if nsteps is None else int(nsteps)
        prm = w.params(nsteps)
        if sor_tol is not None:
            prm.sor_tol = sor_tol
        if sor_max_it is not None:
            prm.sor_max_it = sor_max_it
        t0 = time.perf_counter()
        target = dipole_series(w, theta * 1.03, nsteps, prm)
        t_fwd = time.perf_counter() - t0
        fit = DipoleFit(w, target, nsteps, min(checkpoint_every, nsteps), prm)
        t0 = time.perf_counter()
        loss, grad, _ = fit.loss_grad(theta)
        t_adj = time.perf_counter() - t0
        fds = {}
        t0 = time.perf_counter()
        for h in hs:
            fd = central_difference(fit, theta, h)
            fds[f"{h:g}"] = {"gradient": dict(zip(PARAM_NAMES, fd.tolist())),
                             "relative_difference": dict(zip(PARAM_NAMES, ((grad - fd) / np.abs(fd)).tolist()))}
        t_fd = time.perf_counter() - t0
        fit.close()
        out["horizons"].append({
            "steps": nsteps, "sor": [prm.sor_omega, prm.sor_tol, prm.sor_max_it], "loss": loss,
            "gradient_adjoint": dict(zip(PARAM_NAMES, grad.tolist())), "central_differences": fds,
            "seconds": {"forward_run": t_fwd, "loss_and_adjoint_gradient": t_adj,
                        "central_differences": t_fd}})
        print(json.dumps(out["horizons"][-1]), flush=True)
    return out


if __name__ == "__main__":
    import argparse
    import sys

    ap = argparse.ArgumentParser(description="adjoint gradient vs central differences")
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--steps", default=None, help="comma-separated horizons (default: the workload's steps)")
    ap.add_argument("--checkpoint-every", type=int, default=100)
    ap.add_argument("--h", default="1e-4", help="comma-separated relative finite-difference steps")
    ap.add_argument("--sor-tol", type=float, default=None)
    ap.add_argument("--sor-max-it", type=int, default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    horizons = [int(v) for v in a.steps.split(",")] if a.steps else [None]
    res = study(a.workload, horizons, [float(v) for v in a.h.split(",")], a.checkpoint_every, a.sor_tol,
                a.sor_max_it)
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    sys.exit(0)
