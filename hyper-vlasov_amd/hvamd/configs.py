"""The benchmark workloads of the test-particle loop, as (World, Skyrme,
Params, initial state) builders.

Conventions carried as explicit parameters (the reference code differs from
these workloads in several places):
  - grid half-width d_max = n dx / 2 (the reference derives it from
    d_max_scale * 1.5 A^(1/3), include/init.hpp:329);
  - smearing width from r_fwhm = 1.444 fm (auto_compute.py:54);
  - SOR omega = 1.8 and tolerance 1e-6 on the max-update rule, 512 iterations
    cap (the reference hard-codes 1.5 / 1e-4 / 512, include/physics.hpp:103,
    include/sim_structs.hpp:33,44);
  - boost "protons +dp, neutrons -dp" = HV_BOOST_SYMMETRIC with
    eta = dp / (hbar c) [fm^-1], at t = 0;
  - initial state uniform in a ball of R = 1.2 A^(1/3) and in a Fermi ball of
    k_F = (3 pi^2 rho0 / 2)^(1/3), seeded (include/hv_init.h).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import HV_BOOST_SYMMETRIC
from .sim import HBARC, RHO0, Params, Skyrme, World, sample_uniform_spheres, sigma_from_fwhm

R_FWHM = 1.444


@dataclass
class Workload:
    name: str
    z: int
    n: int
    ppn: int
    grid: int
    dx: float
    dt: float
    steps: int
    dp: float          # boost momentum [MeV/c]
    seed: int = 128
    skyrme: Skyrme = None
    omega: float = 1.8
    tol: float = 1e-6
    max_it: int = 512

    @property
    def A(self) -> int:
        return self.z + self.n

    @property
    def n_protons(self) -> int:
        return self.z * self.ppn

    @property
    def n_neutrons(self) -> int:
        return self.n * self.ppn

    @property
    def total(self) -> int:
        return self.n_protons + self.n_neutrons

    def world(self) -> World:
        d = self.grid * self.dx / 2.0
        return World((self.grid,) * 3, (d,) * 3)

    def params(self, steps: int | None = None) -> Params:
        st = self.steps if steps is None else steps
        return Params(t_f=self.dt * st, steps=st, substeps=1, t_exc=0.0, eta_exc=self.dp / HBARC,
                      boost_mode=HV_BOOST_SYMMETRIC, z=self.z, n=self.n,
                      sigma_r=sigma_from_fwhm(R_FWHM), part_per_nucleon=self.ppn, sor_omega=self.omega,
                      sor_tol=self.tol, sor_max_it=self.max_it)

    def sky(self) -> Skyrme:
        return self.skyrme if self.skyrme is not None else Skyrme(-356.0, 303.0, 32.0, 7.0 / 6.0)

    def initial_state(self):
        """x, y, z, kx, ky, kz (float64, protons first)."""
        r_max = 1.2 * self.A ** (1.0 / 3.0)
        k_f = (3.0 * np.pi ** 2 * RHO0 / 2.0) ** (1.0 / 3.0)
        return sample_uniform_spheres(self.seed, self.n_protons, self.n_neutrons, r_max, k_f)


def c0() -> Workload:
    """Plain FP64, CPU-runnable: A=40, 4,000 tp, 32^3 dx 0.5, dt 0.5, 1,000 steps, dp 20."""
    return Workload("c0_A40_32cubed", 20, 20, 100, 32, 0.5, 0.5, 1000, 20.0)


def c1() -> Workload:
    """Precision form: A=208, 104,000 tp, 64^3 dx 0.4, dt 0.5, 4,000 steps, dp 15."""
    return Workload("c1_A208_64cubed", 82, 126, 500, 64, 0.4, 0.5, 4000, 15.0)


def c2() -> Workload:
    """Throughput: A=208, 1.04M tp, 128^3 dx 0.25, dt 0.25, 2,000 steps, dp 15."""
    return Workload("c2_A208_128cubed", 82, 126, 5000, 128, 0.25, 0.25, 2000, 15.0)


def c4() -> Workload:
    """Differentiation: A=120, 120,000 tp, 64^3 dx 0.4, dt 0.5, 1,500 steps."""
    return Workload("c4_A120_64cubed", 50, 70, 1000, 64, 0.4, 0.5, 1500, 15.0)


def c3(n_instances: int = 256, seed: int = 128) -> list:
    """Ensemble: n instances, A ~ U[16, 238], Z = round(0.4 A), test particles
    per nucleon ~ U[200, 1000], grid 32^3, 64^3 or 96^3 over a box of
    U[16, 24] fm, Skyrme B and C scaled by U[0.9, 1.1], dp ~ U[5, 30] MeV/c,
    1,500 steps of dt = 0.5 fm/c.  Instance i draws from its own seeded
    stream, so the first k instances do not depend on n."""
    out = []
    for i in range(n_instances):
        rng = np.random.default_rng([seed, i])
        A = int(rng.integers(16, 239))
        Z = int(round(0.4 * A))
        ppn = int(rng.integers(200, 1001))
        n = int(rng.choice([32, 64, 96]))
        box = float(rng.uniform(16.0, 24.0))
        fb, fc = rng.uniform(0.9, 1.1, 2)
        dp = float(rng.uniform(5.0, 30.0))
        sk = Skyrme(-356.0, 303.0 * float(fb), 32.0 * float(fc), 7.0 / 6.0)
        out.append(Workload(f"c3_{i}_A{A}_{n}cubed", Z, A - Z, ppn, n, box / n, 0.5, 1500, dp,
                            seed=seed * 1000003 + i, skyrme=sk))
    return out


WORKLOADS = {"c0": c0, "c1": c1, "c2": c2, "c4": c4}
