"""hv_stats_noalloc (the diagnostics without a context or any allocation)
against the context entry points hv_center_of_mass / hv_mean_squared_radius
on the same positions, bit for bit, in both precisions: two calls in a row,
the second with fewer particles (the scratch keeps the first call's partial
rows; the second call must not read them)."""
import ctypes

import numpy as np
import pytest

import hvamd
from hvamd import _lib, configs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """The few HIP runtime calls a caller makes to hold device arrays."""

    def __init__(self):
        _lib.load()
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def download(self, p, a):
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H) == 0
        return a

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _context_stats(w, prec, x, y, z, kx, ky, kz, np_, nn):
    sim = hvamd.Simulation(w.world(), np_, nn, prec)
    try:
        sim.upload_particles(x, y, z, kx, ky, kz)
        return np.array([sim.center_of_mass(0), sim.center_of_mass(1), sim.mean_squared_radius(0),
                         sim.mean_squared_radius(1)])
    finally:
        sim.close()


@pytest.mark.parametrize("prec", [hvamd.HV_PREC_F64, hvamd.HV_PREC_MIXED])
def test_stats_noalloc_equals_context_entry_points(prec):
    w = configs.c1()
    x, y, z, kx, ky, kz = w.initial_state()
    x = x * 1.9  # some particles outside the 12.8 fm grid: the in-box counts matter
    lib = _lib.load()
    hip = _Hip()
    dt = np.float64 if prec == hvamd.HV_PREC_F64 else np.float32
    try:
        out = hip.empty(32)
        world = w.world().c()
        sizes = [(w.n_protons, w.n_neutrons), (w.n_protons // 3, w.n_neutrons // 5)]
        for np_, nn in sizes:
            sel = np.r_[0:np_, w.n_protons:w.n_protons + nn]
            xs, ys, zs = (np.ascontiguousarray(a[sel]) for a in (x, y, z))
            want = _context_stats(w, prec, xs, ys, zs, kx[sel], ky[sel], kz[sel], np_, nn)
            dx, dy, dz = (hip.upload(np.ascontiguousarray(a.astype(dt))) for a in (xs, ys, zs))
            print(f"hv_stats_noalloc prec={prec} np={np_} nn={nn}", flush=True)
            rc = lib.hv_stats_noalloc(ctypes.byref(world), prec, dx, dy, dz, np_, nn, out, None)
            assert rc == 0
            got = hip.download(out, np.zeros(4))
            print("noalloc", got.tolist(), "context", want.tolist())
            assert np.array_equal(got, want)
    finally:
        hip.free()
