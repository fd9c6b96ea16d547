"""The reference's Woods-Saxon / Fermi-level initial state
(initialize_particles, include/physics.hpp:35-97) and its pieces, over the
C-ABI of include/hv_init.h.  Defaults are the reference's (auto_compute.py:
31-54): V0 = -50 MeV, a = 0.66 fm, epsilon_p = -8 MeV, epsilon_n = -12 MeV,
k_fwhm = 0.346 fm^-1, r_fwhm = 1.444 fm; the reference seeds the C library's
rand() with 128 (src/hyper_vlasov.cpp:154)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import HvWoodsSaxon, ptr
from .sim import Skyrme, sigma_from_fwhm


@dataclass
class WoodsSaxon:
    V0: float
    R12: float
    a: float


def _ws_array(ws):
    arr = (HvWoodsSaxon * 2)()
    for i, w in enumerate(ws):
        arr[i] = HvWoodsSaxon(w.V0, w.R12, w.a)
    return arr


def _ws_list(arr):
    return [WoodsSaxon(arr[i].V0, arr[i].R12, arr[i].a) for i in range(2)]


def _check(name, rc):
    if rc < 0:
        raise _lib.HvError(name, rc, "invalid arguments or device error")
    return rc


def nuclear_radius(mass_number: int) -> float:
    return _lib.load().hv_ws_nuclear_radius(mass_number)


def max_particles(r_max: float, k_max: float, ppn: int) -> int:
    return _lib.load().hv_ws_max_particles(r_max, k_max, ppn)


def sample_round(z, n, ppn, sigma_k, ws, fermi):
    """One sampling round (hv_ws_sample_round); the C library's rand() stream
    continues from where it stands.  Returns (particles dict, fermi[2] after,
    delta[2], counts[6])."""
    lib = _lib.load()
    tot = (z + n) * ppn
    out = {k: np.zeros(tot) for k in ("x", "y", "z", "kx", "ky", "kz")}
    f = np.ascontiguousarray(fermi, dtype=np.float64).copy()
    d = np.zeros(2)
    c = np.zeros(6, dtype=np.int32)
    _check("hv_ws_sample_round", lib.hv_ws_sample_round(
        z, n, ppn, sigma_k, _ws_array(ws), f.ctypes.data_as(_lib._DP), d.ctypes.data_as(_lib._DP),
        c.ctypes.data_as(_lib._IP), *[ptr(out[k]) for k in ("x", "y", "z", "kx", "ky", "kz")]))
    return out, f, d, c


def particle_densities(x, y, z, np_, nn, sigma_r, ppn):
    lib = _lib.load()
    rp, rn = np.zeros(np_ + nn), np.zeros(np_ + nn)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
    _check("hv_ws_particle_densities", lib.hv_ws_particle_densities(*[ptr(v) for v in a], np_, nn, sigma_r, ppn,
                                                                   ptr(rp), ptr(rn)))
    return rp, rn


def fit(x, y, z, rho_p, rho_n, np_, nn, skyrme: Skyrme, ws):
    """Returns (fitted wells [2], LM steps per species)."""
    lib = _lib.load()
    arr = _ws_array(ws)
    its = np.zeros(2, dtype=np.int32)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, rho_p, rho_n)]
    _check("hv_ws_fit", lib.hv_ws_fit(*[ptr(v) for v in a], np_, nn, ctypes.byref(skyrme.c()), arr,
                                      its.ctypes.data_as(_lib._IP)))
    return _ws_list(arr), its


def chi_squared(x, y, z, rho_p, rho_n, np_, nn, skyrme: Skyrme, ws, ppn):
    lib = _lib.load()
    out = np.zeros(2)
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, rho_p, rho_n)]
    _check("hv_ws_chi_squared", lib.hv_ws_chi_squared(*[ptr(v) for v in a], np_, nn, ctypes.byref(skyrme.c()),
                                                      _ws_array(ws), ppn, out.ctypes.data_as(_lib._DP)))
    return out


def woods_saxon_nucleus(z: int, n: int, ppn: int, skyrme: Skyrme | None = None, r_fwhm: float = 1.444,
                        k_fwhm: float = 0.346, V0: float = -50.0, a: float = 0.66, epsilon_p: float = -8.0,
                        epsilon_n: float = -12.0, seed: int = 128) -> dict:
    """The reference's initial nucleus (hv_init_woods_saxon; needs a GPU for
    the particle densities).  Returns the particles (x..kz, protons first),
    the wells, the Fermi levels, the rounds run, chi squared and whether the
    levels settled."""
    lib = _lib.load()
    sk = skyrme if skyrme is not None else Skyrme(-356.8, 303.9, 32.0, 7.0 / 6.0)
    tot = (z + n) * ppn
    out = {k: np.zeros(tot) for k in ("x", "y", "z", "kx", "ky", "kz")}
    fermi = np.array([epsilon_p, epsilon_n], dtype=np.float64)
    ws = (HvWoodsSaxon * 2)()
    it = ctypes.c_int()
    chi = np.zeros(2)
    rc = _check("hv_init_woods_saxon", lib.hv_init_woods_saxon(
        z, n, ppn, ctypes.byref(sk.c()), sigma_from_fwhm(r_fwhm), sigma_from_fwhm(k_fwhm), V0, a,
        fermi.ctypes.data_as(_lib._DP), seed, *[ptr(out[k]) for k in ("x", "y", "z", "kx", "ky", "kz")], ws,
        ctypes.byref(it), chi.ctypes.data_as(_lib._DP)))
    out.update(ws=_ws_list(ws), fermi=fermi, iterations=it.value, chi2=chi,
               converged=not (rc & _lib.HV_WARN_INIT_NOT_CONVERGED))
    return out
