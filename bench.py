#!/usr/bin/env python3
"""Throughput of the hyper-vlasov test-particle step loop on one MI355X.

Workload (default): A=208 (Z=82, N=126), 5,000 test particles per nucleon
(1,040,000 tp), 128^3 grid with dx = 0.25 fm, dt = 0.25 fm/c, dp = 15 MeV/c,
SOR omega 1.8 / tol 1e-6 / 512 iterations.  A step is one pass of the loop
(gather + kick + drift + deposit + smearing + Skyrme + SOR Coulomb solve +
diagnostics) over all test particles with inputs resident in HBM.

Prints ONE JSON line.  `value` is the mixed-precision form (FP32 particles and
grid, FP64 Coulomb potential / SOR / reductions); the plain FP64 form is timed
the same way and reported beside it.  Each form runs the workload's whole
2,000 steps (`full_run`); `value` times the `--steps` steps after `--warmup`
untimed ones in the middle of that run.  Progress goes to stderr.
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
    """Roofline entry for the SOR kernel that dominates a run's device time,
    over every launch of the profiled run.

    Algorithmic bytes of one SOR iteration per interior node: read the iterate
    and its source, write the iterate (FP64 potential: 8 + 8 + 8 = 24 B; the
    mixed form's FP32 correction and FP32 residual source: 4 + 4 + 4 = 12 B).
    A per-sweep launch is half an iteration; one launch of the LDS-resident
    FP32 kernel runs a whole correction solve, so its bytes per launch are its
    iterations x 12 B x nodes (the launch keeps its tile of the correction and
    of its source in LDS, so these bytes are not HBM bytes: `traffic` holds
    the counter-measured HBM bytes)."""
    interior = (w.grid - 2) ** 3
    prof, cnt = res["prof"], res["sor"]
    ms, n = prof.get("sor_resident_f32", (0.0, 0))
    if n > 0:
        it = cnt["iters_f32"] / n
        name = {2: "k_sor_resident_f32d_fold", 1: "k_sor_resident_f32o"}.get(res.get("body"), "k_sor_resident_f32")
        # the deep-halo body of a loop step also computes the residual (read phi
        # and rho) and folds the correction into phi (read and write phi) once
        # per launch: 8 + 4 + 8 + 8 B per node
        once = 28.0 * interior if res.get("body") == 2 else 0.0
        return {"kernel": name + " (the mixed form's FP32 SOR correction, one LDS-resident launch per solve)",
                "bytes_per_launch": it * 12.0 * interior + once, "avg_launch_us": 1e3 * ms / n,
                "iterations_per_launch": it, "launches": n,
                "limiter": "neighbour exchange: the round trip of a tile face between CUs (agent-scope stores "
                           "and loads) each iteration; the tile's values stay in LDS"}
    ms, n = prof.get("sor_sweep_f32", (0.0, 0))
    if n > 0:
        return {"kernel": "k_sor_colour_f32 (one red or black sweep of the FP32 SOR correction)",
                "bytes_per_launch": 6.0 * interior, "avg_launch_us": 1e3 * ms / n, "iterations_per_launch": 0.5,
                "launches": res["issued"]["sor_sweep_f32"], "timed_launches": n}
    ms, n = prof["sor_sweep"]
    return {"kernel": "k_sor_colour (one red or black FP64 SOR sweep)", "bytes_per_launch": 12.0 * interior,
            "avg_launch_us": 1e3 * ms / max(n, 1), "iterations_per_launch": 0.5,
            "launches": res["issued"]["sor_sweep"], "timed_launches": n}


def copy_calibration(mib: int):
    """A plain device copy of `mib` MiB (read mib, write mib) by the library's
    copy kernel (k_copy_calibration), run before the workload so that HBM byte
    counters (FETCH_SIZE / WRITE_SIZE) read for it give the counters' scale
    against a known size."""
    from hvamd.sim import copy_bandwidth
    copy_bandwidth(mib, reps=3)


def measured_copy_gbs(mib: int = 1024, reps: int = 5) -> float:
    """HBM bandwidth of a plain device copy on this GPU (read + write bytes
    per second, best of `reps`, the library's copy kernel), for context beside
    the 8 TB/s spec."""
    from hvamd.sim import copy_bandwidth
    return copy_bandwidth(mib, reps)[0]


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


def window_start(total: int, steps: int, warmup: int) -> int:
    """First step of the warm-up + timed window: the middle of the run (the
    SOR iterations per step climb along a run, so its first steps are its
    cheapest), or as late as the window still fits."""
    return max(0, min(total // 2, total - warmup - steps))


def run_gpu(w, prec: str, steps: int, warmup: int, full: bool = True):
    """Pass 1 (kernel timers off except in the tuner's trials): every step of
    the workload's run, timed piece by piece; the particle-sort cadence is
    autotuned on the live run at its start and again just before the headline
    window (tune_single: the trials are steps of the run); the window is
    `warmup` untimed steps and then `steps` timed steps starting at
    window_start (the middle of the run).  Pass 2: the same run again with
    the window's cadence (sorting changes no bit: the same work, checked by
    the SOR iteration counts) and the per-kernel-class timers on over all its
    steps: kernel_ms_per_step and the roofline kernel's launches over the
    whole run, the launches a kernel trace of the run averages over.  With
    full=False only the tuning and the window are run (from step 0)."""
    import hvamd
    from hvamd.autotune import tune_single, tuning_steps
    p = hvamd.HV_PREC_F64 if prec == "fp64" else hvamd.HV_PREC_MIXED
    x, y, z, kx, ky, kz = w.initial_state()
    need = tuning_steps()
    total = max(w.steps, 2 * need + warmup + steps) if full else need + warmup + steps
    first = max(2 * need, window_start(total, steps, warmup)) if full else need
    prm = w.params(total)
    sk = w.sky()

    def new_sim():
        sim = hvamd.Simulation(w.world(), w.n_protons, w.n_neutrons, p)
        sim.upload_particles(x, y, z, kx, ky, kz)
        sim.run_init(sk, prm)
        sim.sync()
        return sim

    def log(msg):
        print(f"[bench] {w.name} {prec}: {msg}", file=sys.stderr, flush=True)

    sim = new_sim()
    its_all, state, tunes = [], None, []
    walls = {"window": 0.0, "rest": 0.0}

    def plain(a, b, key):
        for c0 in range(a, b, 250):  # pieces of <= 250 steps, one progress line each
            c1 = min(b, c0 + 250)
            t0 = time.perf_counter()
            _, it = sim.run_steps(sk, prm, c0, c1 - c0)
            sim.sync()
            dt = time.perf_counter() - t0
            walls[key] += dt
            its_all.append(it)
            log(f"steps {c0}-{c1} {1e3 * dt / (c1 - c0):.3f} ms/step")

    def tune(at):
        from hvamd.autotune import _Steps
        r = _Steps(sim, sk, prm)
        t0 = time.perf_counter()
        best, timings, nxt = tune_single(sim, sk, prm, at, _runner=r)
        sim.sync()
        walls["rest"] += time.perf_counter() - t0
        its_all.extend(r.its)
        tunes.append({"at_step": at, "best": best,
                      "particle_pass_ms_per_step": {str(k): v for k, v in timings["particle_pass"].items()}})
        log(f"sort cadence tuned at step {at}: {best} {timings['particle_pass']}")
        return nxt

    step = tune(0)
    if full:
        plain(step, first - need, "rest")
        step = tune(first - need)
        assert step == first
    plain(first, first + warmup, "rest")
    if prec == "fp64":
        # the state at the start of the timed window, for the CPU baseline
        state = sim.download_particles()
        state["phi"] = sim.download_field(hvamd.HV_FIELD_COULOMB)
    plain(first + warmup, first + warmup + steps, "window")
    plain(first + warmup + steps, total, "rest")
    sort_every = sim.get_option(hvamd.HV_OPT_SORT_EVERY)
    sim.close()
    its_all = np.concatenate(its_all)
    win = slice(first + warmup, first + warmup + steps)
    sim = new_sim()
    sim.set_option(hvamd.HV_OPT_SORT_EVERY, sort_every)
    sim.profile_enable(True)
    _, its2 = sim.run_steps(sk, prm, 0, total)
    sim.sync()
    prof = sim.profile_read()
    issued = sim.profile_issued()
    sor = sim.sor_counts()
    sim.profile_enable(False)
    body = sim.get_option(hvamd.HV_OPT_SOR_F32_OWNED) if prec == "mixed" else None
    sim.close()
    assert np.array_equal(its_all, its2), "the profiled run must do the timed run's work"
    return dict(wall=walls["window"], its=its_all[win], full_wall=walls["window"] + walls["rest"], full_its=its_all,
                total=total, first=first + warmup, prof=prof, issued=issued, sor=sor, state=state, body=body,
                tunes=tunes, sort_every=sort_every)


def kernel_ms_per_step(res) -> dict:
    """Device ms per step of each kernel class over the profiled run.  The
    per-sweep SOR classes are timed one launch in eight with an event pair
    around the dispatch, which inflates a few-microsecond launch; for them
    the sampled time per launch and the launches per step are reported, and
    the SOR stage's time per step is the timed run's wall time per step minus
    the other classes (the sweeps with their launch gaps and read-backs)."""
    out, other = {}, 0.0
    total = res["total"]
    for k, (ms, n) in res["prof"].items():
        if k in ("sor_sweep", "sor_sweep_f32"):
            if n > 0:
                out[k + "_sampled_us_per_launch"] = 1e3 * ms / n
                out[k + "_launches_per_step"] = res["issued"][k] / total
            continue
        out[k] = ms / total
        other += ms / total
    if res["prof"]["sor_resident_f32"][1] == 0:
        out["sor_stage_wall_remainder"] = 1e3 * res["full_wall"] / total - other
    return out


ENS_SORT_CHOICES = (0, 4, 8, 16)
RETUNE_EVERY = 500  # steps between re-tunings of the ensemble's sort cadence


def run_batch(ws, prec: str, steps: int, warmup: int, mid_steps: int, full: bool, graph: bool = True):
    """The ensemble in one batch.  Both forms: init, the autotuner at step 0
    (smearing tile and form, sort cadence; tune_batch, its trials are steps
    of the run), `warmup` untimed steps and `steps` timed steps: the early
    window, the same steps in both forms.  The mixed form with `full` then
    continues over the workload's whole run, re-tuning the sort cadence
    every RETUNE_EVERY steps (retune_batch_sort; the particles' order decays
    along a run), with `mid_steps` timed steps in the middle of the run (the
    mixed form's headline window) and every step outside the tuner's trials
    timed for the whole-run rate; a few profiled steps after that window give
    the stage breakdown and the SOR kernel's time and work counters (the
    mixed form: the ensemble's LDS-resident FP32 launch; FP64: the per-sweep
    launches)."""
    import hvamd
    from hvamd.autotune import retune_batch_sort, tune_batch
    from hvamd.batch import Batch
    p = hvamd.HV_PREC_F64 if prec == "fp64" else hvamd.HV_PREC_MIXED
    prof_steps = 4
    run_len = ws[0].steps if full else 0
    b = Batch([w.world() for w in ws], [w.n_protons for w in ws], [w.n_neutrons for w in ws], p)
    if not graph:
        b.set_option(hvamd.HV_OPT_BATCH_GRAPH, 0)
    for i, w in enumerate(ws):
        b.instance(i).upload_particles(*w.initial_state())
    sks = [w.sky() for w in ws]
    prm = [w.params(max(run_len, 256)) for w in ws]
    b.run_init(sks, prm)

    def log(msg):
        print(f"[bench] c3 ensemble {prec}: {msg}", file=sys.stderr, flush=True)

    best, timings, step = tune_batch(b, sks, prm, 0, steps_per_trial=2, sort_choices=ENS_SORT_CHOICES,
                                     tile_choices=(16, 64), form_choices=(0, 1))
    tunes = [{"at_step": 0, "best": best,
              "stage_ms_per_step": {f"sort{k[0]}_tile{k[1]}_fused{k[2]}": v for k, v in timings.items()}}]
    log(f"tuned at step 0: {best}")
    b.run_steps(sks, prm, step, warmup)
    step += warmup
    t0 = time.perf_counter()
    _, its = b.run_steps(sks, prm, step, steps)
    wall = time.perf_counter() - t0
    early = {"first_step": step, "steps": steps, "ms_per_step": 1e3 * wall / steps,
             "value": sum(w.total for w in ws) * steps / wall,
             "sor_iters_per_step": {"mean": float(its.mean()), "max": int(its.max())}}
    log(f"steps {step}-{step + steps} {1e3 * wall / steps:.2f} ms/step")
    step += steps
    counted = {"steps": steps, "wall": wall}
    mid = None
    if full:
        mid_first = run_len // 2

        def plain(a, z):
            for c0 in range(a, z, 100):
                c1 = min(z, c0 + 100)
                t = time.perf_counter()
                _, it = b.run_steps(sks, prm, c0, c1 - c0)
                dt = time.perf_counter() - t
                counted["steps"] += c1 - c0
                counted["wall"] += dt
                log(f"steps {c0}-{c1} {1e3 * dt / (c1 - c0):.2f} ms/step, SOR iterations {it.mean():.1f}")

        def retune():
            nonlocal step
            t = step
            best_s, part, step = retune_batch_sort(b, sks, prm, step, sort_choices=ENS_SORT_CHOICES)
            tunes.append({"at_step": t, "best": {"sort_every": best_s},
                          "particle_stage_ms_per_step": {str(k): v for k, v in part.items()}})
            log(f"sort cadence re-tuned at step {t}: {best_s} {part}")

        for at in range(RETUNE_EVERY, run_len, RETUNE_EVERY):
            if at > mid_first and step <= mid_first:
                plain(step, mid_first - 80)
                step = mid_first - 80
                retune()
                plain(step, mid_first)
                step = mid_first
                t = time.perf_counter()
                _, its_m = b.run_steps(sks, prm, step, mid_steps)
                dtm = time.perf_counter() - t
                counted["steps"] += mid_steps
                counted["wall"] += dtm
                mid = {"first_step": step, "steps": mid_steps, "ms_per_step": 1e3 * dtm / mid_steps,
                       "value": sum(w.total for w in ws) * mid_steps / dtm,
                       "sor_iters_per_step": {"mean": float(its_m.mean()), "max": int(its_m.max())}}
                log(f"mid-run window steps {step}-{step + mid_steps} {1e3 * dtm / mid_steps:.2f} ms/step")
                step += mid_steps
                # the stage breakdown and the SOR kernel right after the window
                b.profile_enable(True)
                b.run_steps(sks, prm, step, prof_steps)
                prof = b.profile_read()
                cnt = b.sor_counts()
                b.profile_enable(False)
                step += prof_steps
            if step < at:
                plain(step, at)
                step = at
            retune()
        if step < run_len:
            plain(step, run_len)
            step = run_len
    if mid is None:
        b.profile_enable(True)
        b.run_steps(sks, prm, step, prof_steps)
        prof = b.profile_read()
        cnt = b.sor_counts()
        b.profile_enable(False)
    body = b.get_option(hvamd.HV_OPT_SOR_F32_OWNED) if prec == "mixed" else None
    steered = b.get_option(hvamd.HV_OPT_BATCH_DEVICE_STEER)
    graphs = b.graph_info()
    b.close()
    n_tot = sum(w.total for w in ws)
    node_iters = cnt["node_iters_f32"] if prec == "mixed" else cnt["node_iters_f64"]
    bpn = 12.0 if prec == "mixed" else 24.0
    if prec == "mixed" and prof["sor_resident_f32"][1]:
        # one launch per step; its time is measured for every launch
        ms, nl = prof["sor_resident_f32"]
        launches = nl
        avg_ms = ms / nl
        # the deep-halo body also computes each instance's residual and folds
        # its correction (28 B per interior node once per launch, as for c2)
        once = 28.0 * sum((w.grid - 2) ** 3 for w in ws) * nl if body == 2 else 0.0
        achieved = (node_iters * bpn + once) / (ms * 1e-3) / 1e9
        kname = {2: "kb_sor_resident_f32d (FUSE: with its residual and fold)",
                 1: "kb_sor_resident_f32o"}.get(body, "kb_sor_resident_f32") + \
            " (the FP32 correction of every instance, one LDS-resident launch per step)"
    else:
        kcls = "sor_sweep_f32" if prec == "mixed" else "sor_sweep"
        ms, nl = prof[kcls]
        launches = cnt["sweeps_f32"] if prec == "mixed" else cnt["sweeps_f64"]
        avg_ms = ms / max(nl, 1)
        achieved = node_iters * bpn / (launches * avg_ms * 1e-3) / 1e9 if launches else 0.0
        once = 0.0
        kname = ("kb_sor_colour_f32" if prec == "mixed" else "kb_sor_colour") + \
            " (one red or black sweep over every instance still iterating)"
    head = mid if mid is not None else early
    return {
        "value": head["value"], "ms_per_step": head["ms_per_step"],
        "window": "mid-run" if mid is not None else "early",
        "early_window": early, "mid_run_window": mid,
        "full_run": ({"run_steps": run_len, "steps_timed": counted["steps"],
                      "ms_per_step": 1e3 * counted["wall"] / counted["steps"],
                      "value": n_tot * counted["steps"] / counted["wall"],
                      "note": "every step of the run outside the tuner's trials and the profiled steps"}
                     if full else None),
        "test_particles": n_tot, "instances": len(ws), "grid_nodes": sum(w.grid ** 3 for w in ws),
        "device_steered_step": bool(steered), "graph_steps": graphs["graph_steps"],
        "tunes": tunes,
        "stage_ms_per_step": {k: v[0] / prof_steps for k, v in prof.items() if v[1] and k in
                              ("particle_pass", "smear", "sor_stage", "potential")},
        "stage_steps": f"{prof_steps} profiled steps after the {'mid-run' if mid is not None else 'early'} window",
        "roofline": sweep_or_resident_entry(prec, kname, achieved, avg_ms, launches,
                                            node_iters * bpn + (once if prec == "mixed" else 0.0)),
    }


def sweep_or_resident_entry(prec, kname, achieved, avg_ms, launches, alg_bytes):
    """The ensemble's SOR kernel against the HBM peak.  The FP64 per-sweep
    launches stream the potential through HBM, so their algorithmic rate is
    an HBM rate.  The mixed form's LDS-resident launch keeps every tile in LDS
    and moves only tile faces through memory: its algorithmic rate (12 B per
    node and iteration over the launch time) is what a streaming kernel would
    need, not HBM traffic, and can exceed the HBM peak."""
    e = {"kernel": kname, "peak": HBM_PEAK_GBS, "unit": "GB/s", "avg_launch_us": 1e3 * avg_ms,
         "launches": launches, "bytes_per_launch": alg_bytes / max(launches, 1)}
    if prec == "mixed" and kname.startswith("kb_sor_resident"):
        e.update({"bound": "neighbour exchange (LDS-resident tiles)", "algorithmic_equiv_GBs": achieved,
                  "algorithmic_equiv_over_hbm_peak": achieved / HBM_PEAK_GBS})
    else:
        e.update({"bound": "hbm", "achieved": achieved, "frac": achieved / HBM_PEAK_GBS})
    return e


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(w, state, window_steps: int, window_its, budget_s: float):
    """The CPU restatement (oracle/liboracle.so, OpenMP) over the GPU's timed
    window: from the FP64 run's state at the window's first step, as many of
    the window's steps as fit in about `budget_s` seconds, on this host's
    cores.  The oracle's FP64 loop equals the GPU's FP64 loop bit for bit, so
    these steps run the same SOR iterations as the GPU's window (checked)."""
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
        return time.perf_counter() - t0, its[1:]

    t1, _ = run(1)
    nsteps = int(max(2, min(window_steps, budget_s / max(t1, 1e-3))))
    t, its = run(nsteps)
    same = bool(np.array_equal(its, np.asarray(window_its[:nsteps])))
    return {
        "value": w.total * nsteps / t,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "ms_per_step": 1e3 * t / nsteps,
        "sor_iters_per_step": float(its.mean()),
        "us_per_sor_iteration_and_step": 1e6 * t / nsteps / float(its.mean()),
        "same_sor_iterations_as_gpu_window": same,
        "sample": (f"the first {nsteps} steps of the GPU's timed {w.name} window, from the FP64 GPU run's state "
                   f"there (mean {its.mean():.0f} SOR iterations/step, the GPU's FP64 iterations "
                   f"{'exactly' if same else 'NOT'}), oracle/liboracle.so, "
                   f"OpenMP {cores} threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                   f"OMP_PLACES={os.environ.get('OMP_PLACES')}), -O3 -march=x86-64-v3, {cpu_model()}"),
    }


def reference_literal_c0(budget_s: float):
    """The reference's own step loop (oracle/_ref/libhvref.so: cpu_simulate's
    sequence of the reference's kernels, src/hyper_vlasov.cpp:53-104,
    including the O(G^2) smearing of include/physics.hpp:209-229 and the
    SOR constants it hard-codes, omega 1.5 / tol 1e-4) on the A=40, 32^3
    workload's initial state, OpenMP on this host's cores: per-step time from
    the difference of two run lengths (the pre-loop solve cancels), and the
    oracle's literal and factorised smearing timed on the same counts."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, Reference
    if not Reference.available:
        return None
    from hvamd import configs
    w = configs.c0()
    x, y, z, kx, ky, kz = w.initial_state()
    st = dict(x=x, y=y, z=z, kx=kx, ky=ky, kz=kz)
    ref = Reference()
    prm = w.params()
    sk = w.sky()
    n, d = w.world().n, w.world().d_max

    def run(steps):
        t0 = time.perf_counter()
        ref.simulate(st, w.n_protons, w.n_neutrons, n, d, sk.A, sk.B, sk.C, sk.gamma, prm.dt * steps, steps, 0.0,
                     prm.eta_exc, w.z, w.n, prm.sigma_r, prm.part_per_nucleon)
        return time.perf_counter() - t0

    t1 = run(1)
    extra = int(max(1, min(20, budget_s / max(t1, 1e-3) - 1)))
    t2 = run(1 + extra)
    per_step = max(t2 - t1, 1e-9) / extra
    o = Oracle()
    counts, _, _ = o.deposit(x, y, z, w.n_protons, w.n_neutrons, n, d)
    t0 = time.perf_counter()
    o.smear(counts, n, d, prm.sigma_r, prm.part_per_nucleon, literal=True)
    t_lit = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(10):
        o.smear(counts, n, d, prm.sigma_r, prm.part_per_nucleon)
    t_fac = (time.perf_counter() - t0) / 10
    o.L.hvo_threads.restype = ctypes.c_int
    return {"value": w.total / per_step, "unit": "particle-steps/s", "kind": "reference",
            "cores": o.L.hvo_threads(), "ms_per_step": 1e3 * per_step, "steps_timed": extra,
            "smear_literal_over_factorised": t_lit / t_fac,
            "sample": (f"{extra} steps of the reference's own loop on the {w.name} state (its hard-coded SOR "
                       f"omega 1.5 / tol 1e-4, O(G^2) smearing), oracle/_ref/libhvref.so built from the "
                       f"reference sources, OpenMP; smearing of one 32^3 count field: literal O(G^2) "
                       f"{1e3 * t_lit:.1f} ms vs factorised {1e3 * t_fac:.2f} ms")}


def counter_traffic(rl, workload):
    """`traffic` (HBM bytes per launch from the committed counter profile of
    this kernel) and its fraction of the HBM peak over the launch time."""
    t = measured_traffic(rl, workload)
    if t is None:
        return None, None
    return t, t / (rl["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c2", choices=["c0", "c1", "c2", "c4"])
    ap.add_argument("--full", type=int, default=1,
                    help="run every step of the workload (the timed window sits mid-run); 0: tuning and the window only")
    ap.add_argument("--ensemble", type=int, default=256, help="instances of the c3 ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=8, help="timed steps of the early ensemble window")
    ap.add_argument("--ensemble-mid-steps", type=int, default=40,
                    help="timed steps of the mixed ensemble's mid-run window (with --full 1)")
    ap.add_argument("--fp64", type=int, default=1, help="also time the plain FP64 form")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain MIB-MiB device copy (HBM counter scale)")
    ap.add_argument("--single", type=int, default=1,
                    help="0: the ensemble only (a profiling run; its line carries no headline value)")
    ap.add_argument("--ensemble-graph", type=int, default=1,
                    help="0: launch the ensemble's steps one kernel at a time instead of replaying step graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build measures the loop on a "
              "single MI355X and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's OpenMP placement (read when libgomp starts)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import hvamd
    from hvamd import configs
    hvamd.load()
    w = configs.WORKLOADS[args.workload]()
    copy_gbs = measured_copy_gbs()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    full = bool(args.full)
    if not args.single:
        ens = configs.c3(args.ensemble)
        g = bool(args.ensemble_graph)
        bm = run_batch(ens, "mixed", args.ensemble_steps, 2, args.ensemble_mid_steps, full, g)
        b64 = run_batch(ens, "fp64", args.ensemble_steps, 2, 0, False, g) if args.fp64 else None
        print(json.dumps({"ensemble_c3": {"mixed": bm, "fp64": b64}, "measured_copy_GBs": copy_gbs}), flush=True)
        return
    mixed = run_gpu(w, "mixed", args.steps, args.warmup, full)
    fp64 = run_gpu(w, "fp64", args.steps, args.warmup, full) if args.fp64 else None

    value = w.total * args.steps / mixed["wall"]
    rl = sor_roofline(w, mixed)
    achieved = rl["bytes_per_launch"] / (rl["avg_launch_us"] * 1e-6) / 1e9
    traffic, traffic_frac = counter_traffic(rl, args.workload)
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
                   "dt_fm_c": w.dt, "sor": [w.omega, w.tol, w.max_it], "window_first_step": mixed["first"],
                   "run_steps": mixed["total"], "parallelism": "single GPU"},
        "sor_iters_per_step": float(np.mean(mixed["its"])),
        "sor_counts": mixed["sor"],
        "kernel_ms_per_step": kernel_ms_per_step(mixed),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_frac_of_peak": traffic_frac,
                     **rl},
    }
    out["sort_cadence_autotune"] = {"mixed": mixed["tunes"], "fp64": fp64["tunes"] if fp64 else None}
    out["full_run"] = {"mixed": {"value": w.total * mixed["total"] / mixed["full_wall"],
                                 "ms_per_step": 1e3 * mixed["full_wall"] / mixed["total"], "steps": mixed["total"],
                                 "sor_iters_per_step": float(mixed["full_its"].mean())}}
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
                       "kernel_ms_per_step": kernel_ms_per_step(fp64),
                       "sor_achieved_GBs": r64["bytes_per_launch"] / (r64["avg_launch_us"] * 1e-6) / 1e9,
                       "sor_avg_launch_us": r64["avg_launch_us"], "sor_launches": r64["launches"],
                       "sor_timed_launches": r64["timed_launches"]}
        out["mixed_over_fp64"] = value / v64
        out["full_run"]["fp64"] = {"value": w.total * fp64["total"] / fp64["full_wall"],
                                   "ms_per_step": 1e3 * fp64["full_wall"] / fp64["total"], "steps": fp64["total"],
                                   "sor_iters_per_step": float(fp64["full_its"].mean())}
        out["full_run"]["mixed_over_fp64"] = out["full_run"]["mixed"]["value"] / out["full_run"]["fp64"]["value"]
        if args.cpu_budget > 0:
            out["cpu_baseline"] = cpu_baseline(w, fp64["state"], args.steps, fp64["its"], args.cpu_budget)
            lit = reference_literal_c0(args.cpu_budget / 3)
            if lit is not None:
                out["cpu_baseline_reference_loop_c0"] = lit
    if args.ensemble > 0:
        ens = configs.c3(args.ensemble)
        g = bool(args.ensemble_graph)
        bm = run_batch(ens, "mixed", args.ensemble_steps, 2, args.ensemble_mid_steps, full, g)
        b64 = run_batch(ens, "fp64", args.ensemble_steps, 2, 0, False, g) if args.fp64 else None
        out["ensemble_c3"] = {
            "mixed": bm, "fp64": b64,
            # like for like: the same early steps in both forms (FP64 runs only those)
            "mixed_over_fp64_early_window": bm["early_window"]["value"] / b64["value"] if b64 else None,
            "mixed_over_single_run_fp64": bm["value"] / out["fp64"]["value"] if args.fp64 else None}
    out["roofline"]["measured_copy_GBs"] = copy_gbs
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
