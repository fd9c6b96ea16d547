"""Gradient of a dipole-response misfit with respect to the Skyrme
coefficients (hv_dipole_loss_grad: the discrete adjoint of the FP64 loop),
wrapped as a torch.autograd.Function for fitting scripts; WindowedDipoleFit
sums the misfit over windows of a long horizon, each started from the
target trajectory's state (multiple shooting).

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


class WindowedDipoleFit:
    """The misfit over a long horizon as a sum over windows (multiple
    shooting): window w covers steps [s_w, s_w + n_w) and starts from the
    state of the target trajectory at s_w (its particles, fields rebuilt by
    hv_run_init), so each window's derivative is a short-horizon derivative
    of the discrete loop, and
        L = sum_w sum_{s in w} (D_w(t_s) - D*(t_s))^2,   dL/dtheta = sum_w dL_w/dtheta.
    Over a single window this is DipoleFit's loss.  The boost of the run is
    part of the target's state after step 0, so windows after the first
    start without one.  target_states[w]: the particle arrays (x, y, z, kx,
    ky, kz) at s_w, from target_run."""

    def __init__(self, workload, target, target_states, window: int, checkpoint_every: int = 100, params=None):
        self.w = workload
        self.target = np.ascontiguousarray(target, dtype=np.float64)
        self.nsteps = len(self.target)
        self.window = int(window)
        self.ck = int(checkpoint_every)
        self.params = params if params is not None else workload.params(self.nsteps)
        self.windows = [(s0, min(self.window, self.nsteps - s0)) for s0 in range(0, self.nsteps, self.window)]
        if len(target_states) != len(self.windows):
            raise ValueError("one target state per window")
        self.states = target_states
        self.sim = Simulation(workload.world(), workload.n_protons, workload.n_neutrons, HV_PREC_F64)
        self.evaluations = 0

    def close(self):
        self.sim.close()

    def _params(self, s0, n):
        import copy
        p = copy.copy(self.params)
        dt = self.params.t_f / self.params.steps
        p.steps = n
        p.t_f = dt * n
        if s0 > 0:
            p.t_exc = 1e30  # the boost is already in the window's starting state
        return p

    def loss_grad(self, theta, grad: bool = True):
        """(loss, gradient[4] or None, D series) at theta = (A, B, C, gamma)."""
        sk = Skyrme(*[float(v) for v in theta]).c()
        total = 0.0
        g = np.zeros(4) if grad else None
        d = np.zeros(self.nsteps)
        dp = _lib._DP
        for (s0, n), st in zip(self.windows, self.states):
            self.sim.upload_particles(st["x"], st["y"], st["z"], st["kx"], st["ky"], st["kz"])
            pr = self._params(s0, n).c()
            loss = ctypes.c_double()
            dw = np.zeros(n)
            gw = np.zeros(4) if grad else None
            tw = np.ascontiguousarray(self.target[s0:s0 + n])
            rc = self.sim.lib.hv_dipole_loss_grad(self.sim.ctx, ctypes.byref(sk), ctypes.byref(pr), n,
                                                  tw.ctypes.data_as(dp), min(self.ck, n), ctypes.byref(loss),
                                                  dw.ctypes.data_as(dp), gw.ctypes.data_as(dp) if grad else None)
            self.sim._rc("hv_dipole_loss_grad", rc)
            total += loss.value
            d[s0:s0 + n] = dw
            if grad:
                g += gw
        self.evaluations += 1
        return total, g, d


def target_run(workload, theta, nsteps: int, window: int, params=None):
    """(D series, particle states at the window starts 0, window, 2 window, ...)
    of one run at theta (the target of a windowed fit)."""
    from .sim import Simulation as _Sim
    prm = params if params is not None else workload.params(nsteps)
    sim = _Sim(workload.world(), workload.n_protons, workload.n_neutrons, HV_PREC_F64)
    try:
        sim.upload_particles(*workload.initial_state())
        sk = Skyrme(*[float(v) for v in theta])
        sim.run_init(sk, prm)
        states, stats = [], []
        for s0 in range(0, nsteps, window):
            n = min(window, nsteps - s0)
            states.append(sim.download_particles())
            st, _ = sim.run_steps(sk, prm, s0, n)
            stats.append(st)
        stats = np.concatenate(stats)
        dfac = workload.n * workload.z / (workload.n + workload.z)
        return dfac * (stats[:, 0] - stats[:, 1]), states
    finally:
        sim.close()


def dipole_series(workload, theta, nsteps: int, params=None) -> np.ndarray:
    """D(t_s) of a run at theta (e.g. the target of a fit)."""
    fit = DipoleFit(workload, np.zeros(nsteps), nsteps, max(nsteps, 1), params)
    try:
        return fit.loss_grad(theta, grad=False)[2]
    finally:
        fit.close()


def central_difference(fit: DipoleFit, theta, h: float = 1e-4, log=None) -> np.ndarray:
    """Central finite differences with relative steps h |theta_k| (log: called
    with a line after each component)."""
    theta = np.asarray(theta, dtype=np.float64)
    out = np.zeros(4)
    for k in range(4):
        step = h * abs(theta[k])
        tp, tm = theta.copy(), theta.copy()
        tp[k] += step
        tm[k] -= step
        out[k] = (fit.loss_grad(tp, grad=False)[0] - fit.loss_grad(tm, grad=False)[0]) / (2.0 * step)
        if log:
            log(f"h {h:g} {PARAM_NAMES[k]}: central difference {out[k]:.10g}")
    return out


try:
    # the library (and the ROCm runtime it is built against) before torch, so
    # that the process holds one HIP runtime
    try:
        _lib.load()
    except Exception:
        pass
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


def study_windowed(name: str = "c4", nsteps: int | None = None, window: int = 100, hs=(1e-5,),
                   checkpoint_every: int = 100, sor_tol: float | None = None,
                   sor_max_it: int | None = None) -> dict:
    """The workload's whole horizon as a windowed misfit (WindowedDipoleFit):
    the target run at theta * 1.03 with its states at the window starts, the
    adjoint gradient of the summed misfit at theta, and central differences
    of the same summed misfit at each relative step h."""
    import time

    from . import configs
    w = configs.WORKLOADS[name]()
    nsteps = w.steps if nsteps is None else int(nsteps)
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    prm = w.params(nsteps)
    if sor_tol is not None:
        prm.sor_tol = sor_tol
    if sor_max_it is not None:
        prm.sor_max_it = sor_max_it
    import sys

    def log(msg):
        print(f"[adjoint] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    target, states = target_run(w, theta * 1.03, nsteps, window, prm)
    t_tgt = time.perf_counter() - t0
    log(f"target run: {nsteps} steps, {len(states)} window states, {t_tgt:.1f} s")
    fit = WindowedDipoleFit(w, target, states, window, checkpoint_every, prm)
    try:
        t0 = time.perf_counter()
        loss, grad, _ = fit.loss_grad(theta)
        t_adj = time.perf_counter() - t0
        log(f"loss {loss:.10g}, adjoint gradient {grad.tolist()}, {t_adj:.1f} s")
        fds = {}
        t0 = time.perf_counter()
        for h in hs:
            fd = central_difference(fit, theta, h, log)
            fds[f"{h:g}"] = {"gradient": dict(zip(PARAM_NAMES, fd.tolist())),
                             "relative_difference": dict(zip(PARAM_NAMES, ((grad - fd) / np.abs(fd)).tolist()))}
        t_fd = time.perf_counter() - t0
    finally:
        fit.close()
    out = {"workload": w.name, "test_particles": w.total, "steps": nsteps, "window": window,
           "windows": len(states), "checkpoint_every": checkpoint_every,
           "sor": [prm.sor_omega, prm.sor_tol, prm.sor_max_it], "loss": loss,
           "gradient_adjoint": dict(zip(PARAM_NAMES, grad.tolist())), "central_differences": fds,
           "seconds": {"target_run": t_tgt, "loss_and_adjoint_gradient": t_adj, "central_differences": t_fd}}
    print(json.dumps(out), flush=True)
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
    ap.add_argument("--window", type=int, default=0,
                    help="with a window length: the horizon as a windowed misfit (multiple shooting)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    horizons = [int(v) for v in a.steps.split(",")] if a.steps else [None]
    if a.window > 0:
        res = [study_windowed(a.workload, hz, a.window, [float(v) for v in a.h.split(",")], a.checkpoint_every,
                              a.sor_tol, a.sor_max_it) for hz in horizons]
    else:
        res = study(a.workload, horizons, [float(v) for v in a.h.split(",")], a.checkpoint_every, a.sor_tol,
                    a.sor_max_it)
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    sys.exit(0)
