"""Host-side mirror of the reference's step-loop interface over libhvamd.so.

`Simulation` owns one device context.  Its methods carry the reference's
function names (include/physics.hpp, include/tools.hpp) and the same in-place
meaning, so a loop written against the reference reads the same here:

    sim.distribute_volumetric_particles_cic()      # include/tools.hpp:36-96
    sim.compute_volumetric_densities(sigma_r, ppn) # include/physics.hpp:203-234
    sim.compute_volumetric_skyrme_potentials(skm)  # include/physics.hpp:273-281
    sim.compute_volumetric_coulomb_potentials_sor(omega, tol, max_it)
    sim.merge_volumetric_potentials()              # include/tools.hpp:378-384
    sim.compute_volumetric_forces_fdm()            # include/physics.hpp:160-200
    sim.distribute_forces_to_particles_cic()       # include/tools.hpp:99-168
    sim.update_momenta_half(dt); sim.update_positions_full(dt)
    sim.nuclear_excitation(eta, z, n, mode)
    sim.center_of_mass(type); sim.mean_squared_radius(type)

Errors: the reference exit(1)s on allocation failure and prints a warning on
SOR non-convergence; here a negative status raises HvError and warnings are
returned as the HV_WARN_* bit mask (and kept in `last_warnings`).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (HV_BOOST_REF_WEIGHTED, HV_BOOST_SYMMETRIC, HV_FIELD_COULOMB,
                   HV_FIELD_DENSITY, HV_FIELD_FORCES_X, HV_FIELD_FORCES_Y, HV_FIELD_FORCES_Z,
                   HV_FIELD_POTENTIALS, HV_NEUTRONS, HV_PREC_F64, HV_PREC_MIXED, HV_PROTONS,
                   HvError, HvParams, HvSkyrme, HvWorld, ptr)

MC2 = 935.0
HBARC = 197.33
RHO0 = 0.16


def sigma_from_fwhm(fwhm: float) -> float:
    """calc_sigma (include/physics_formulas.hpp:58-62)."""
    return fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))


@dataclass
class World:
    """World<T> (include/sim_structs.hpp:73-76)."""
    n: tuple = (32, 32, 32)
    d_max: tuple = (8.0, 8.0, 8.0)

    @property
    def G(self) -> int:
        return int(self.n[0] * self.n[1] * self.n[2])

    def c(self) -> HvWorld:
        return HvWorld((ctypes.c_int * 3)(*self.n), (ctypes.c_double * 3)(*self.d_max))


@dataclass
class Skyrme:
    """Skyrme<T> (include/sim_structs.hpp:84-86)."""
    A: float = -356.0
    B: float = 303.0
    C: float = 32.0
    gamma: float = 7.0 / 6.0

    def c(self) -> HvSkyrme:
        return HvSkyrme(self.A, self.B, self.C, self.gamma)


@dataclass
class Params:
    """Loop parameters (Parameters<T>, include/sim_structs.hpp:49-53, plus the
    SOR knobs the reference hard-codes and the boost convention)."""
    t_f: float = 500.0
    steps: int = 1000
    substeps: int = 1
    t_exc: float = 0.0
    eta_exc: float = 0.0
    boost_mode: int = HV_BOOST_SYMMETRIC
    z: int = 20
    n: int = 20
    sigma_r: float = field(default_factory=lambda: sigma_from_fwhm(1.444))
    part_per_nucleon: int = 100
    sor_omega: float = 1.8
    sor_tol: float = 1e-6
    sor_max_it: int = 512

    @property
    def dt(self) -> float:
        return self.t_f / self.steps

    def c(self) -> HvParams:
        return HvParams(self.t_f, self.steps, self.substeps, self.t_exc, self.eta_exc,
                        self.boost_mode, self.z, self.n, self.sigma_r, self.part_per_nucleon,
                        self.sor_omega, self.sor_tol, self.sor_max_it)


def _check(fn: str, rc: int, ctx=None) -> int:
    if rc < 0:
        lib = _lib.load()
        msg = lib.hv_last_error(ctx).decode() if ctx else ""
        raise HvError(fn, rc, msg)
    return rc


def sample_uniform_spheres(seed: int, n_protons: int, n_neutrons: int, r_max: float, k_max: float):
    """Seeded initial state (include/hv_init.h): returns x, y, z, kx, ky, kz."""
    lib = _lib.load()
    tot = n_protons + n_neutrons
    arrs = [np.empty(tot) for _ in range(6)]
    _check("hv_sample_uniform_spheres",
           lib.hv_sample_uniform_spheres(seed, n_protons, n_neutrons, r_max, k_max, *[ptr(a) for a in arrs]))
    return arrs


class Simulation:
    """One device context: particles protons-first, fields on the world grid."""

    def __init__(self, world: World, n_protons: int, n_neutrons: int, precision: int = HV_PREC_F64):
        self.lib = _lib.load()
        self.world = world
        self.np = int(n_protons)
        self.nn = int(n_neutrons)
        self.total = self.np + self.nn
        self.precision = precision
        self._w = world.c()
        h = ctypes.c_void_p()
        rc = self.lib.hv_ctx_create(ctypes.byref(self._w), self.np, self.nn, precision, ctypes.byref(h))
        if rc < 0:
            raise HvError("hv_ctx_create", rc, "bad world/particle counts or device allocation failure")
        self.ctx = h
        self.owned = True
        self.last_warnings = 0

    @classmethod
    def view(cls, ctx, world: World, n_protons: int, n_neutrons: int, precision: int) -> "Simulation":
        """A non-owning wrapper of a context that something else owns (an
        instance of a Batch)."""
        self = cls.__new__(cls)
        self.lib = _lib.load()
        self.world = world
        self.np, self.nn = int(n_protons), int(n_neutrons)
        self.total = self.np + self.nn
        self.precision = precision
        self._w = world.c()
        self.ctx = ctypes.c_void_p(ctx)
        self.owned = False
        self.last_warnings = 0
        return self

    # -- lifetime ---------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            if self.owned:
                self.lib.hv_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _rc(self, fn: str, rc: int) -> int:
        _check(fn, rc, self.ctx)
        self.last_warnings = rc
        return rc

    # -- transfer ---------------------------------------------------------------
    def upload_particles(self, x=None, y=None, z=None, kx=None, ky=None, kz=None, fx=None, fy=None, fz=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (x, y, z, kx, ky, kz, fx, fy, fz)]
        for a in arrs:
            if a is not None and a.shape != (self.total,):
                raise ValueError(f"particle arrays must have shape ({self.total},)")
        return self._rc("hv_upload_particles", self.lib.hv_upload_particles(self.ctx, *[ptr(a) for a in arrs]))

    def download_particles(self):
        """Returns dict of x, y, z, kx, ky, kz, fx, fy, fz (float64)."""
        names = ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")
        arrs = [np.empty(self.total) for _ in names]
        self._rc("hv_download_particles", self.lib.hv_download_particles(self.ctx, *[ptr(a) for a in arrs]))
        return dict(zip(names, arrs))

    def _field_len(self, which: int) -> int:
        return self.world.G if which == HV_FIELD_COULOMB else 2 * self.world.G

    def upload_field(self, which: int, values):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if a.size != self._field_len(which):
            raise ValueError("field size mismatch")
        return self._rc("hv_upload_field", self.lib.hv_upload_field(self.ctx, which, ptr(a)))

    def download_field(self, which: int) -> np.ndarray:
        a = np.empty(self._field_len(which))
        self._rc("hv_download_field", self.lib.hv_download_field(self.ctx, which, ptr(a)))
        return a

    # -- the reference's kernels ------------------------------------------------
    def distribute_volumetric_particles_cic(self) -> int:
        return self._rc("hv_cic_deposit", self.lib.hv_cic_deposit(self.ctx))

    def compute_volumetric_densities(self, sigma_r: float, part_per_nucleon: int) -> int:
        return self._rc("hv_smear_density", self.lib.hv_smear_density(self.ctx, sigma_r, part_per_nucleon))

    def compute_volumetric_skyrme_potentials(self, skyrme: Skyrme) -> int:
        s = skyrme.c()
        return self._rc("hv_skyrme_potentials", self.lib.hv_skyrme_potentials(self.ctx, ctypes.byref(s)))

    def set_initial_coulomb_boundaries(self, z: int) -> int:
        return self._rc("hv_set_coulomb_bc", self.lib.hv_set_coulomb_bc(self.ctx, z))

    def compute_volumetric_coulomb_potentials_sor(self, omega=1.5, tol=1e-4, max_it=512):
        """Returns (iterations, last max |dphi|)."""
        it = ctypes.c_int()
        md = ctypes.c_double()
        self._rc("hv_coulomb_sor", self.lib.hv_coulomb_sor(self.ctx, omega, tol, max_it, ctypes.byref(it),
                                                            ctypes.byref(md)))
        return it.value, md.value

    def merge_volumetric_potentials(self) -> int:
        return self._rc("hv_merge_potentials", self.lib.hv_merge_potentials(self.ctx))

    def compute_volumetric_forces_fdm(self) -> int:
        return self._rc("hv_forces_fdm", self.lib.hv_forces_fdm(self.ctx))

    def distribute_forces_to_particles_cic(self) -> int:
        return self._rc("hv_cic_gather", self.lib.hv_cic_gather(self.ctx))

    def update_momenta_half(self, dt: float) -> int:
        return self._rc("hv_kick_half", self.lib.hv_kick_half(self.ctx, dt))

    def update_positions_full(self, dt: float) -> int:
        return self._rc("hv_drift", self.lib.hv_drift(self.ctx, dt))

    def nuclear_excitation(self, eta: float, z: int, n: int, mode: int = HV_BOOST_REF_WEIGHTED) -> int:
        return self._rc("hv_excite", self.lib.hv_excite(self.ctx, eta, z, n, mode))

    def total_energy(self, skyrme: Skyrme, part_per_nucleon: int) -> np.ndarray:
        """hv_total_energy: [E, E_kin, E_sky, E_coul] of the current state [MeV]."""
        out = np.zeros(4)
        s = skyrme.c()
        self._rc("hv_total_energy", self.lib.hv_total_energy(self.ctx, ctypes.byref(s), part_per_nucleon,
                                                             out.ctypes.data_as(_lib._DP)))
        return out

    def center_of_mass(self, type_: int) -> float:
        out = ctypes.c_double()
        self._rc("hv_center_of_mass", self.lib.hv_center_of_mass(self.ctx, type_, ctypes.byref(out)))
        return out.value

    def mean_squared_radius(self, type_: int) -> float:
        out = ctypes.c_double()
        self._rc("hv_mean_squared_radius", self.lib.hv_mean_squared_radius(self.ctx, type_, ctypes.byref(out)))
        return out.value

    # -- fused loop ---------------------------------------------------------------
    def run_init(self, skyrme: Skyrme, params: Params) -> int:
        s, p = skyrme.c(), params.c()
        it = ctypes.c_int()
        self._rc("hv_run_init", self.lib.hv_run_init(self.ctx, ctypes.byref(s), ctypes.byref(p), ctypes.byref(it)))
        return it.value

    def run_steps(self, skyrme: Skyrme, params: Params, first_step: int, nsteps: int):
        """Returns (stats[nsteps, 4], sor_iters[nsteps])."""
        s, p = skyrme.c(), params.c()
        stats = np.zeros((nsteps, 4))
        its = np.zeros(nsteps, dtype=np.int32)
        self._rc("hv_run_steps", self.lib.hv_run_steps(self.ctx, ctypes.byref(s), ctypes.byref(p), first_step,
                                                       nsteps, ptr(stats), ptr(its)))
        return stats, its

    def simulate(self, skyrme: Skyrme, params: Params):
        """hv_simulate: pre-loop + all steps.  Returns (stats[steps, 4], sor_iters[steps])."""
        s, p = skyrme.c(), params.c()
        stats = np.zeros((params.steps, 4))
        its = np.zeros(params.steps, dtype=np.int32)
        self._rc("hv_simulate", self.lib.hv_simulate(self.ctx, ctypes.byref(s), ctypes.byref(p), ptr(stats),
                                                     ptr(its)))
        return stats, its

    # -- timing -----------------------------------------------------------------
    def timer_start(self):
        self._rc("hv_timer_start", self.lib.hv_timer_start(self.ctx))

    def timer_stop(self) -> float:
        ms = ctypes.c_double()
        self._rc("hv_timer_stop", self.lib.hv_timer_stop(self.ctx, ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._rc("hv_sync", self.lib.hv_sync(self.ctx))

    def set_option(self, option: int, value: int):
        self._rc("hv_set_option", self.lib.hv_set_option(self.ctx, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        self._rc("hv_get_option", self.lib.hv_get_option(self.ctx, option, ctypes.byref(v)))
        return v.value

    def profile_enable(self, on: bool = True):
        self._rc("hv_profile_enable", self.lib.hv_profile_enable(self.ctx, 1 if on else 0))

    def profile_read(self) -> dict:
        """{class name: (device ms, launches)} accumulated since profile_enable."""
        out = {}
        for k, name in enumerate(_lib.KCLASS_NAMES):
            ms = ctypes.c_double()
            n = ctypes.c_longlong()
            self._rc("hv_profile_read", self.lib.hv_profile_read(self.ctx, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def profile_issued(self) -> dict:
        """{class name: launches issued since profile_enable}: for the sampled
        SOR sweep classes every launch, timed or not."""
        out = {}
        for k, name in enumerate(_lib.KCLASS_NAMES):
            n = ctypes.c_longlong()
            self._rc("hv_profile_issued", self.lib.hv_profile_issued(self.ctx, k, ctypes.byref(n)))
            out[name] = n.value
        return out

    def sor_counts(self) -> dict:
        """SOR solves, FP32 correction iterations and FP64 iterations since
        profile_enable."""
        v = [ctypes.c_longlong() for _ in range(3)]
        self._rc("hv_sor_counts", self.lib.hv_sor_counts(self.ctx, *[ctypes.byref(x) for x in v]))
        return {"solves": v[0].value, "iters_f32": v[1].value, "iters_f64": v[2].value}


def copy_bandwidth(mib: int = 1024, reps: int = 5) -> tuple:
    """(GB/s, microseconds) of the best of `reps` plain device copies of
    `mib` MiB by the library's copy kernel (read + write bytes per second)."""
    lib = _lib.load()
    gbs, us = ctypes.c_double(), ctypes.c_double()
    rc = lib.hv_copy_bandwidth(mib << 20, reps, ctypes.byref(gbs), ctypes.byref(us))
    if rc < 0:
        raise _lib.HvError("hv_copy_bandwidth", rc, "copy probe failed")
    return gbs.value, us.value


def dipole(stats: np.ndarray, z: int, n: int) -> np.ndarray:
    """D(t) = N Z / A (<z>_p - <z>_n)  (scripts/programs.py:99)."""
    return (n * z) / (n + z) * (stats[:, 0] - stats[:, 1])
