"""Precision study of the mixed FP32/FP64 form against the plain FP64 form.

Runs one workload twice on the GPU (FP64, whose loop equals the CPU
restatement bit for bit, and mixed), plus an FP64 run whose initial positions
are moved by one FP32 rounding step (relative 2^-24), which measures how far
the dynamics itself amplifies a perturbation of that size (the noise floor a
precision change cannot beat).  Reported:
  - D(t) = N Z / A (<z>_p - <z>_n) (scripts/programs.py:99): max and rms
    |D_x(t) - D_64(t)| relative to max |D_64|;
  - total energy (hv_total_energy) every `chunk` steps: drift
    (E(T) - E(0)) / |E(0)| of each run and max |E_x - E_64| / |E_64(0)|;
  - dipole-response centroid E_c (5-40 MeV window) of each run and its
    difference from FP64, over the whole run and over its first `window`
    steps;
  - the same D(t) error over growing horizons, and the first step at which it
    exceeds 1e-4 of the amplitude;
  - each of these for several perturbed FP64 runs (noise in different
    directions), as a distribution, and where the mixed form falls in it.

The dynamics of these workloads is chaotic: a relative perturbation of 2^-24
in the positions grows past 1e-4 of the D(t) amplitude within a few hundred
steps, so agreement is judged against that noise floor, horizon by horizon.

    python -m hvamd.precision --workload c1 [--steps N] [--chunk 100] [--out file.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, spectrum
from ._lib import HV_OPT_SOR_EXTRAPOLATE, HV_PREC_F64, HV_PREC_MIXED
from .sim import HBARC, Simulation


def perturbed_positions(w, seed: int, size: float = 2.0 ** -24):
    """The workload's initial positions, each coordinate scaled by 1 + s size
    with s = +1 for seed 0 (every coordinate moved outward) and otherwise a
    seeded random sign per particle and coordinate: one FP32 rounding step of
    position noise, in a different direction for every seed."""
    x, y, z, kx, ky, kz = w.initial_state()
    if seed == 0:
        sx = sy = sz = 1.0
    else:
        rng = np.random.default_rng([seed, 2 ** 24])
        sx, sy, sz = (rng.choice([-1.0, 1.0], size=x.size) for _ in range(3))
    return x * (1.0 + sx * size), y * (1.0 + sy * size), z * (1.0 + sz * size), kx, ky, kz


def run(w, prec: int, steps: int, chunk: int, perturb_seed: int | None = None, extrapolate: int | None = None):
    """(stats[steps, 4], energies[(step, [E, Ekin, Esky, Ecoul])], sor iterations, seconds);
    extrapolate: the mixed form's HV_OPT_SOR_EXTRAPOLATE (None: the default)."""
    if perturb_seed is None:
        x, y, z, kx, ky, kz = w.initial_state()
    else:
        x, y, z, kx, ky, kz = perturbed_positions(w, perturb_seed)
    sim = Simulation(w.world(), w.n_protons, w.n_neutrons, prec)
    try:
        if extrapolate is not None and prec == HV_PREC_MIXED:
            sim.set_option(HV_OPT_SOR_EXTRAPOLATE, extrapolate)
        sim.upload_particles(x, y, z, kx, ky, kz)
        prm = w.params(steps)
        sk = w.sky()
        t0 = time.perf_counter()
        sim.run_init(sk, prm)
        energies = [(0, sim.total_energy(sk, w.ppn).tolist())]
        stats, its = [], []
        for first in range(0, steps, chunk):
            c = min(chunk, steps - first)
            s, it = sim.run_steps(sk, prm, first, c)
            stats.append(s)
            its.append(it)
            energies.append((first + c, sim.total_energy(sk, w.ppn).tolist()))
        sim.sync()
        return np.concatenate(stats), energies, np.concatenate(its), time.perf_counter() - t0
    finally:
        sim.close()


def compare(w, ref, other, window: int) -> dict:
    d_ref = spectrum.dipole_moment(ref[0][:, 0], ref[0][:, 1], w.z, w.n)
    d_oth = spectrum.dipole_moment(other[0][:, 0], other[0][:, 1], w.z, w.n)
    amp = np.abs(d_ref).max()
    diff = np.abs(d_oth - d_ref)
    e_ref = np.array([e for _, e in ref[1]])
    e_oth = np.array([e for _, e in other[1]])
    eta = w.dp / HBARC
    c_ref = spectrum.dipole_response(ref[0], w.dt, w.z, w.n, eta)[4]
    c_oth = spectrum.dipole_response(other[0], w.dt, w.z, w.n, eta)[4]
    cw_ref = spectrum.dipole_response(ref[0][:window], w.dt, w.z, w.n, eta)[4]
    cw_oth = spectrum.dipole_response(other[0][:window], w.dt, w.z, w.n, eta)[4]
    over = diff / amp > 1e-4
    horizons = {}
    for h in (50, 100, 250, 500, 1000, 2000, 4000):
        if h <= len(diff):
            horizons[str(h)] = float(diff[:h].max() / np.abs(d_ref[:h]).max())
    return {
        "D_max_abs_diff_over_amplitude": float(diff.max() / amp),
        "D_rms_diff_over_amplitude": float(np.sqrt((diff ** 2).mean()) / amp),
        "D_max_rel_diff_by_horizon": horizons,
        "D_first_step_rel_diff_above_1e-4": int(np.argmax(over)) if over.any() else None,
        "energy_drift": float((e_oth[-1, 0] - e_oth[0, 0]) / abs(e_oth[0, 0])),
        "energy_max_abs_diff_over_E0": float(np.abs(e_oth[:, 0] - e_ref[:, 0]).max() / abs(e_ref[0, 0])),
        "centroid_MeV": float(c_oth),
        "centroid_diff_MeV": float(c_oth - c_ref),
        f"centroid_first_{window}_steps_MeV": float(cw_oth),
        f"centroid_first_{window}_steps_diff_MeV": float(cw_oth - cw_ref),
    }


def distribution(values) -> dict:
    v = np.asarray([x for x in values if x is not None], dtype=np.float64)
    if v.size == 0:
        return {"n": 0}
    return {"n": int(v.size), "min": float(v.min()), "median": float(np.median(v)), "max": float(v.max()),
            "values": v.tolist()}


def study(name: str = "c1", steps: int | None = None, chunk: int = 100, window: int = 1000,
          n_perturbed: int = 6) -> dict:
    """FP64, mixed and n_perturbed FP64 runs with 2^-24 position noise in
    different directions (seeds 0..n-1): the mixed form's errors against FP64
    beside the distribution of the perturbed runs' errors (the chaos floor)."""
    w = configs.WORKLOADS[name]()
    steps = w.steps if steps is None else steps
    window = min(window, steps)
    f64 = run(w, HV_PREC_F64, steps, chunk)
    print(f"[precision] {w.name}: FP64 done ({f64[3]:.1f} s)", file=sys.stderr, flush=True)
    mix = run(w, HV_PREC_MIXED, steps, chunk)
    print(f"[precision] {w.name}: mixed done ({mix[3]:.1f} s)", file=sys.stderr, flush=True)
    mix0 = run(w, HV_PREC_MIXED, steps, chunk, extrapolate=0)
    print(f"[precision] {w.name}: mixed, no extrapolation, done ({mix0[3]:.1f} s)", file=sys.stderr, flush=True)
    pert = []
    for seed in range(n_perturbed):
        r = run(w, HV_PREC_F64, steps, chunk, perturb_seed=seed)
        pert.append(compare(w, f64, r, window))
        print(f"[precision] {w.name}: perturbed seed {seed} done ({r[3]:.1f} s)", file=sys.stderr, flush=True)
    e64 = np.array([e for _, e in f64[1]])
    d64 = spectrum.dipole_moment(f64[0][:, 0], f64[0][:, 1], w.z, w.n)
    mixed = compare(w, f64, mix, window)
    floor = {}
    for key in ("D_max_abs_diff_over_amplitude", "D_rms_diff_over_amplitude", "D_first_step_rel_diff_above_1e-4",
                "energy_max_abs_diff_over_E0", "centroid_diff_MeV", f"centroid_first_{window}_steps_diff_MeV"):
        floor[key] = distribution([r[key] for r in pert])
    floor["abs_centroid_diff_MeV"] = distribution([abs(r["centroid_diff_MeV"]) for r in pert])
    floor[f"abs_centroid_first_{window}_steps_diff_MeV"] = distribution(
        [abs(r[f"centroid_first_{window}_steps_diff_MeV"]) for r in pert])
    floor["D_max_rel_diff_by_horizon"] = {h: distribution([r["D_max_rel_diff_by_horizon"][h] for r in pert])
                                          for h in mixed["D_max_rel_diff_by_horizon"]}

    def rank(v, vals):
        # perturbed runs whose value is at most the mixed form's, of all of them
        return f"{sum(1 for x in vals if x is not None and abs(x) <= abs(v))}/{len(vals)}"

    position = {
        "abs_centroid_diff": rank(mixed["centroid_diff_MeV"], [r["centroid_diff_MeV"] for r in pert]),
        f"abs_centroid_first_{window}_steps_diff": rank(
            mixed[f"centroid_first_{window}_steps_diff_MeV"],
            [r[f"centroid_first_{window}_steps_diff_MeV"] for r in pert]),
        "D_max_rel_diff_by_horizon": {h: rank(v, [r["D_max_rel_diff_by_horizon"][h] for r in pert])
                                      for h, v in mixed["D_max_rel_diff_by_horizon"].items()},
    }
    return {
        "workload": w.name, "steps": steps, "dt": w.dt, "test_particles": w.total, "energy_every": chunk,
        "fp64": {"seconds": f64[3], "sor_iters_mean": float(f64[2].mean()),
                 "energy_drift": float((e64[-1, 0] - e64[0, 0]) / abs(e64[0, 0])),
                 "E0_MeV": float(e64[0, 0]), "D_amplitude": float(np.abs(d64).max()),
                 "energy_series_MeV": [[s_, e[0]] for s_, e in f64[1]],
                 "centroid_MeV": float(spectrum.dipole_response(f64[0], w.dt, w.z, w.n, w.dp / HBARC)[4])},
        "mixed_vs_fp64": {"seconds": mix[3], "sor_iters_mean": float(mix[2].mean()), **mixed},
        "mixed_no_extrapolation_vs_fp64": {"seconds": mix0[3], "sor_iters_mean": float(mix0[2].mean()),
                                           **compare(w, f64, mix0, window)},
        "fp64_perturbed_2^-24_vs_fp64": {"runs": pert, "distribution": floor},
        "mixed_within_perturbed": {"note": "k/n: perturbed runs whose |difference| is at most the mixed form's",
                                   **position},
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description="mixed vs FP64 precision study")
    ap.add_argument("--workload", default="c1", choices=sorted(configs.WORKLOADS))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--window", type=int, default=1000, help="steps of the early-window centroid")
    ap.add_argument("--perturbed", type=int, default=6, help="perturbed FP64 runs (different noise directions)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    res = study(a.workload, a.steps, a.chunk, a.window, a.perturbed)
    txt = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    print(txt, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
