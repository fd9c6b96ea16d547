"""The autograd wrapper with theta on the GPU: the loss and its gradient come
back on theta's device.  torch here binds the ROCm runtime libhvamd.so loaded
first (hvamd preloads it), so the process holds one HIP runtime."""
import numpy as np
import pytest

from hvamd import configs
from hvamd.adjoint import DipoleFit, DipoleLoss, dipole_series

pytestmark = pytest.mark.gpu


def test_dipole_loss_on_gpu_theta():
    torch = pytest.importorskip("torch")
    import os
    maps = open(f"/proc/{os.getpid()}/maps").read()
    hip = sorted({ln.split()[-1] for ln in maps.splitlines() if "libamdhip64" in ln})
    print("HIP runtime mapped:", hip, flush=True)
    assert len(hip) == 1
    assert torch.cuda.is_available()
    w = configs.c0()
    nsteps = 12
    prm = w.params(nsteps)
    prm.sor_tol = 1e-11
    prm.sor_max_it = 4000
    sk = w.sky()
    theta = np.array([sk.A, sk.B, sk.C, sk.gamma])
    target = dipole_series(w, theta * 1.03, nsteps, prm)
    fit = DipoleFit(w, target, nsteps, 6, prm)
    try:
        th = torch.tensor(theta, dtype=torch.float64, device="cuda:0", requires_grad=True)
        loss = DipoleLoss.apply(th, fit)
        assert loss.device == th.device
        (3.0 * loss).backward()
        assert th.grad.device == th.device
        _, grad, _ = fit.loss_grad(theta)
    finally:
        fit.close()
    assert torch.equal(th.grad.cpu(), torch.as_tensor(3.0 * grad))
