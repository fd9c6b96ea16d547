#!/usr/bin/env python3
"""Throughput of the hyper-vlasov test-particle step loop on one MI355X.

Workload (default): A=208 (Z=82, N=126), 5,000 test particles per nucleon
(1,040,000 tp), 128^3 grid with dx = 0.25 fm, dt = 0.25 fm/c, dp = 15 MeV/c,
SOR omega 1.8 / tol 1e-6 / 512 iterations.  A step is one pass of the loop
(gather + kick + drift + deposit + smearing + Skyrme + SOR Coulomb solve +
diagnostics) over all test particles with inputs resident in HBM.

Prints ONE JSON line.  `value` is the mixed-precision form (FP32 particles and
grid, FP64 Coulomb potential / SOR / reductions); the plain FP64 form is timed
the same way and reported beside it.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hyper-vlasov_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
FP32_PEAK_TFLOPS = 157.3  # MI355X vector FP32 (counts packed FP32, which this build does not use)
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_c2.json"


def measured_traffic(rl, workload):
    """HBM bytes per launch of the roofline kernel from the committed counter
    profile (FETCH_SIZE / WRITE_SIZE per dispatch, scaled by a plain copy of
    known size), for the 128^3 workload only; None when not profiled."""
    if workload != "c2" or not TRAFFIC_FILE.exists():
        return None
    d = json.loads(TRAFFIC_FILE.read_text())
    if rl["kernel"].startswith("k_sor_resident_f32"):
        # the counters hold for the body they were taken on (its kernel name)
        profiled = d.get("k_sor_resident_f32_kernel", "").split("<")[0]
        if profiled != rl["kernel"].split(" ")[0]:
            return None
        return d["k_sor_resident_f32_bytes_per_iteration"] * rl["iterations_per_launch"]
    for name, v in d["kernels"].items():
        if rl["kernel"].startswith("k_sor_colour_f32") and name.startswith("k_sor_colour_f32"):
            return v["hbm_bytes"]
        if rl["kernel"].startswith("k_sor_colour (") and name == "k_sor_colour<double>":
            return v["hbm_bytes"]
    return None


def sor_roofline(w, res):
    """Roofline entry for the SOR kernel that dominates a run's device time.

    Algorithmic bytes of one SOR iteration per interior node: read the iterate
    and its source, write the iterate (FP64 potential: 8 + 8 + 8 = 24 B; the
    mixed form's FP32 correction and FP32 residual source: 4 + 4 + 4 = 12 B).
    A per-sweep launch is half an iteration; one launch of the LDS-resident
    FP32 kernel runs a whole correction solve (its iteration count is read
    back per solve), so its bytes per launch are iterations x 12 B x nodes."""
    interior = (w.grid - 2) ** 3
    prof, cnt = res["prof"], res["sor"]
    ms, n = prof.get("sor_resident_f32", (0.0, 0))
    if n > 0:
        it = cnt["iters_f32"] / n
        name = {2: "k_sor_resident_f32d", 1: "k_sor_resident_f32o"}.get(res.get("body"), "k_sor_resident_f32")
        return {"kernel": name + " (the mixed form's FP32 SOR correction, one LDS-resident launch per solve)",
                "bytes_per_launch": it * 12.0 * interior, "avg_launch_us": 1e3 * ms / n,
                "iterations_per_launch": it, "launches": n}
    ms, n = prof.get("sor_sweep_f32", (0.0, 0))
    if n > 0:
        return {"kernel": "k_sor_colour_f32 (one red or black sweep of the FP32 SOR correction)",
                "bytes_per_launch": 6.0 * interior, "avg_launch_us": 1e3 * ms / n, "iterations_per_launch": 0.5,
                "launches": n}
    ms, n = prof["sor_sweep"]
    return {"kernel": "k_sor_colour (one red or black FP64 SOR sweep)", "bytes_per_launch": 12.0 * interior,
            "avg_launch_us": 1e3 * ms / max(n, 1), "iterations_per_launch": 0.5, "launches": n}


def copy_calibration(mib: int):
    """A plain device copy of `mib` MiB (read mib, write mib), run three times
    before the workload so that HBM byte counters (FETCH_SIZE / WRITE_SIZE)
    read for it give the counter's scale against a known size."""
    import torch
    a = torch.ones(mib << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def measured_copy_gbs(mib: int = 1024, reps: int = 5) -> float:
    """HBM bandwidth of a plain device copy on this GPU (read + write bytes
    per second, best of `reps`), for context beside the 8 TB/s spec."""
    import torch
    a = torch.ones(mib << 20, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    best = 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = max(best, 2.0 * (mib << 20) / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del a, b
    torch.cuda.synchronize()
    return best


def smear_flops_per_call(w) -> float:
    """Algorithmic floating-point operations of one smearing of the
    workload's grid: per node and species, 4 per z level (two products, two
    sums), 2 per (a, b) tap of the y factor, 2 per x term and the final
    product, over the spherical tap table of hv_smear_table_build (radius
    4 sigma_r; sigma_r from the reference's r_fwhm = 1.444 fm)."""
    sigma = 1.444 / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    cut2 = 16.0 * sigma * sigma
    n = w.grid
    h = w.world().d_max[0] * 2.0 / n  # node spacing (2 d_max / n)
    r = 0
    while r + 1 < n and ((r + 1) * h) ** 2 <= cut2:
        r += 1
    taps_y = 0
    for a in range(r + 1):
        kb = max(b for b in range(r + 1) if (a * h) ** 2 + (b * h) ** 2 <= cut2)
        taps_y += 2 * kb + 1
    per_node = 4 * r + 2 * taps_y + 2 * (2 * r + 1) + 1
    return 2.0 * n ** 3 * per_node


def run_gpu(w, prec: str, steps: int, warmup: int):
    """warmup steps and `steps` timed steps with no kernel timers; then the
    same run again (bit for bit the same work) with the per-kernel-class
    timers on over the same `steps` steps: kernel_ms_per_step and the SOR
    kernel's launches for the roofline."""
    import hvamd
    p = hvamd.HV_PREC_F64 if prec == "fp64" else hvamd.HV_PREC_MIXED
    x, y, z, kx, ky, kz = w.initial_state()
    prm = w.params(warmup + steps)
    sk = w.sky()

    def start():
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, p)
        sim.upload_particles(x, y, z, kx, ky, kz)
        sim.run_init(sk, prm)
        if warmup:
            sim.run_steps(sk, prm, 0, warmup)
        sim.sync()
        return sim

    sim = start()
    t0 = time.perf_counter()
    stats, its = sim.run_steps(sk, prm, warmup, steps)
    sim.sync()
    wall = time.perf_counter() - t0
    state = None
    if prec == "fp64":
        state = sim.download_particles()
        state["phi"] = sim.download_field(hvamd.HV_FIELD_COULOMB)
    sim.close()
    sim = start()
    sim.profile_enable(True)
    _, its2 = sim.run_steps(sk, prm, warmup, steps)
    sim.sync()
    prof = sim.profile_read()
    sor = sim.sor_counts()
    sim.profile_enable(False)
    body = sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED) if prec == "mixed" else None
    sim.close()
    assert np.array_equal(its, its2), "the profiled replay must do the timed run's work"
    return dict(wall=wall, its=its, prof=prof, sor=sor, stats=stats, state=state, body=body)


def run_batch(ws, prec: str, steps: int, tune: bool):
    """The ensemble in one batch: init, autotuned warm-up (smearing tiles and
    form, sort cadence), `steps` timed steps, then a few profiled steps for the
    SOR kernel's time and work counters (the mixed form: the ensemble's
    LDS-resident FP32 launch; FP64: the per-sweep launches)."""
    import hvamd
    from hvamd.autotune import tune_batch
    from hvamd.batch import Batch
    p = hvamd.HV_PREC_F64 if prec == "fp64" else hvamd.HV_PREC_MIXED
    prof_steps = 4
    total_steps = 16 + steps + prof_steps
    b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], p)
    for i, w in enumerate(ws):
        b.instance(i).upload_particles(*w.initial_state())
    sks = [w.sky() for w in ws]
    prm = [w.params(total_steps + 32) for w in ws]
    b.run_init(sks, prm)
    step = 0
    best, timings = None, None
    if tune:
        best, timings, step = tune_batch(b, sks, prm, 0, steps_per_trial=1, sort_choices=(0, 8),
                                         tile_choices=(16, 64), form_choices=(0, 1))
    b.run_steps(sks, prm, step, 2)
    step += 2
    t0 = time.perf_counter()
    _, its = b.run_steps(sks, prm, step, steps)
    wall = time.perf_counter() - t0
    step += steps
    b.profile_enable(True)
    b.run_steps(sks, prm, step, prof_steps)
    prof = b.profile_read()
    cnt = b.sor_counts()
    b.profile_enable(False)
    body = b.get_option(hvamd.HV_OPT_SOR_F32_OWNED) if prec == "mixed" else None
    b.close()
    n_tot = sum(w.total for w in ws)
    node_iters = cnt["node_iters_f32"] if prec == "mixed" else cnt["node_iters_f64"]
    bpn = 12.0 if prec == "mixed" else 24.0
    if prec == "mixed" and prof["sor_resident_f32"][1]:
        # one launch per step; its time is measured for every launch
        ms, nl = prof["sor_resident_f32"]
        launches = nl
        avg_ms = ms / nl
        achieved = node_iters * bpn / (ms * 1e-3) / 1e9
        kname = {2: "kb_sor_resident_f32d", 1: "kb_sor_resident_f32o"}.get(body, "kb_sor_resident_f32") + \
            " (the FP32 correction of every instance, one LDS-resident launch per step)"
    else:
        kcls = "sor_sweep_f32" if prec == "mixed" else "sor_sweep"
        ms, nl = prof[kcls]
        launches = cnt["sweeps_f32"] if prec == "mixed" else cnt["sweeps_f64"]
        avg_ms = ms / max(nl, 1)
        achieved = node_iters * bpn / (launches * avg_ms * 1e-3) / 1e9 if launches else 0.0
        kname = ("kb_sor_colour_f32" if prec == "mixed" else "kb_sor_colour") + \
            " (one red or black sweep over every instance still iterating)"
    return {
        "value": n_tot * steps / wall, "ms_per_step": 1e3 * wall / steps, "test_particles": n_tot,
        "instances": len(ws), "grid_nodes": sum(w.grid ** 3 for w in ws),
        "sor_iters_per_step": {"mean": float(its.mean()), "max": int(its.max())},
        "tuned": best, "tuning_ms_per_step": {f"sort{k[0]}_tile{k[1]}_fused{k[2]}": v
                                              for k, v in (timings or {}).items()},
        "stage_ms_per_step": {k: v[0] / prof_steps for k, v in prof.items() if v[1] and k in
                              ("particle_pass", "smear", "sor_stage", "potential")},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "avg_launch_us": 1e3 * avg_ms, "launches": launches,
                     "bytes_per_launch": node_iters * bpn / max(launches, 1)},
    }


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(w, state, budget_s: float):
    """The CPU restatement (oracle/liboracle.so, OpenMP) from the same warm
    mid-run state, on this host's cores; a bounded number of steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import OParams, OSkyrme, Oracle, oworld, p
    o = Oracle()
    L = o.L
    L.hvo_threads.restype = ctypes.c_int
    cores = L.hvo_threads()  # the OpenMP team the oracle's loop runs on
    n, d = w.world().n, w.world().d_max
    prm = w.params()
    sk = w.sky()
    world = oworld(n, d)
    s = OSkyrme(sk.A, sk.B, sk.C, sk.gamma)

    def run(nsteps):
        st = {k: state[k].copy() for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")}
        phi = state["phi"].copy()
        op = OParams(prm.dt, nsteps, 1e30, 0.0, prm.boost_mode, w.z, w.n, prm.sigma_r, prm.part_per_nucleon,
                     prm.sor_omega, prm.sor_tol, prm.sor_max_it)
        stats = np.zeros(4 * nsteps)
        its = np.zeros(nsteps + 1, dtype=np.int32)
        t0 = time.perf_counter()
        L.hvo_simulate_loop(*[p(st[k]) for k in ("x", "y", "z", "kx", "ky", "kz", "fx", "fy", "fz")],
                            w.n_protons, w.n_neutrons, ctypes.byref(world), ctypes.byref(s), ctypes.byref(op),
                            p(phi), p(stats), p(its))
        return time.perf_counter() - t0, its

    t1, _ = run(1)
    nsteps = int(max(2, min(40, budget_s / max(t1, 1e-3))))
    t, its = run(nsteps)
    return {
        "value": w.total * nsteps / t,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{nsteps} steps of the {w.name} loop from the GPU's warm mid-run state "
                   f"(mean {its[1:].mean():.0f} SOR iterations/step), oracle/liboracle.so, "
                   f"OpenMP {cores} threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                   f"OMP_PLACES={os.environ.get('OMP_PLACES')}), -O3 -march=x86-64-v3, {cpu_model()}"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c4"])
    ap.add_argument("--ensemble", type=int, default=256, help="instances of the c3 ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=8)
    ap.add_argument("--fp64", type=int, default=1, help="also time the plain FP64 form")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain MIB-MiB device copy (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the single-GPU number is the "
              "figure of record for this loop", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's OpenMP placement (read when libgomp starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from hvamd import configs
    w = configs.WORKLOADS[args.workload]()
    # torch's HIP runtime measures the copy before the library's own runtime
    # starts in this process (the two do not share the device afterwards)
    copy_gbs = measured_copy_gbs()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    mixed = run_gpu(w, "mixed", args.steps, args.warmup)
    fp64 = run_gpu(w, "fp64", args.steps, args.warmup) if args.fp64 else None

    value = w.total * args.steps / mixed["wall"]
    rl = sor_roofline(w, mixed)
    achieved = rl["bytes_per_launch"] / (rl["avg_launch_us"] * 1e-6) / 1e9
    out = {
        "metric": "particle-steps/s at 1.04M tp, 128^3 (mixed FP32/FP64 form; plain FP64 beside it)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded uniform-sphere nucleus)",
        "config": {"workload": w.name, "test_particles": w.total, "grid": w.grid, "dx_fm": w.dx,
                   "dt_fm_c": w.dt, "sor": [w.omega, w.tol, w.max_it], "parallelism": "single GPU"},
        "sor_iters_per_step": float(np.mean(mixed["its"])),
        "sor_counts": mixed["sor"],
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in mixed["prof"].items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(rl, args.workload), **rl},
    }
    sm_ms = mixed["prof"]["smear"][0] / max(mixed["prof"]["smear"][1], 1)
    sflop = smear_flops_per_call(w)
    out["smear_flops"] = {"flop_per_call": sflop, "avg_call_us": 1e3 * sm_ms,
                          "achieved_TFLOPs": sflop / (sm_ms * 1e-3) / 1e12, "peak_TFLOPs_fp32": FP32_PEAK_TFLOPS,
                          "frac": sflop / (sm_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS}
    if fp64 is not None:
        v64 = w.total * args.steps / fp64["wall"]
        r64 = sor_roofline(w, fp64)
        out["fp64"] = {"value": v64, "ms_per_step": 1e3 * fp64["wall"] / args.steps,
                       "sor_iters_per_step": float(np.mean(fp64["its"])),
                       "kernel_ms_per_step": {k: v[0] / args.steps for k, v in fp64["prof"].items()},
                       "sor_achieved_GBs": r64["bytes_per_launch"] / (r64["avg_launch_us"] * 1e-6) / 1e9,
                       "sor_avg_launch_us": r64["avg_launch_us"]}
        out["mixed_over_fp64"] = value / v64
        if args.cpu_budget > 0:
            out["cpu_baseline"] = cpu_baseline(w, fp64["state"], args.cpu_budget)
    if args.ensemble > 0:
        ens = configs.c3(args.ensemble)
        bm = run_batch(ens, "mixed", args.ensemble_steps, tune=True)
        b64 = run_batch(ens, "fp64", args.ensemble_steps, tune=True) if args.fp64 else None
        out["ensemble_c3"] = {"mixed": bm, "fp64": b64,
                              "mixed_over_fp64": bm["value"] / b64["value"] if b64 else None,
                              "mixed_over_single_run_fp64": bm["value"] / out["fp64"]["value"] if args.fp64 else None}
    out["roofline"]["measured_copy_GBs"] = copy_gbs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
