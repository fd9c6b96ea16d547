"""Autotuner of the particle-sort cadence and the smearing tiles and form, for
an ensemble (tune_batch) and for a single run (tune_single, and
run_retuned, which re-tunes the sort cadence along a run as the particles'
order decays).

No knob changes a bit of the results (sorting only reorders particles in
storage; the deposits are exact integer sums; the smearing tile width and form
only regroup the same operations), so they are tuned on the live ensemble:
each candidate advances the run by a few steps and is timed, and the fastest
is left set (compared on the stages the knobs act on, not on whole steps).
Smearing tiles narrower than the widest that fits the LDS trade
halo re-reads for more blocks; the fused form keeps the D channels in LDS but
walks the planes of a tile in one block, the two-pass form stores them and
runs every (plane, tile) as a block of its own; sorting costs a radix sort per
instance but makes the deposit's LDS boxes compact."""
from __future__ import annotations

import itertools

from ._lib import HV_OPT_SMEAR_FUSED, HV_OPT_SMEAR_TILE, HV_OPT_SORT_EVERY

SORT_CHOICES = (0, 8, 32)
TILE_CHOICES = (16, 32, 64)
FORM_CHOICES = (0, 1)


def _stage_ms(batch, stages, skyrmes, params, step, nsteps):
    """Milliseconds per step of the named stages over nsteps steps (stage
    timers on: the stages are timed alone, the SOR's varying work stays out of
    the comparison)."""
    batch.profile_enable(True)
    batch.run_steps(skyrmes, params, step, nsteps)
    prof = batch.profile_read()
    batch.profile_enable(False)
    return sum(prof[k][0] for k in stages) / nsteps


def tune_batch(batch, skyrmes, params, first_step: int, steps_per_trial: int = 2, sort_choices=SORT_CHOICES,
               tile_choices=TILE_CHOICES, form_choices=FORM_CHOICES):
    """Returns (best {sort_every, smear_tile, smear_fused}, timings
    {(sort, tile, fused): ms/step of the stages the knobs act on}, next
    first_step).

    The SOR work per step changes along a run, so the candidates are compared
    on the stages their knobs act on only: the smearing tile and form on the
    smearing stage, the sort cadence on the particle stage with its sorts
    (over max(sort_every, steps_per_trial) steps, so that the sort's cost is
    included once per cadence).  Each setting first runs one untimed step (its
    one-time costs: tile tables, a first sort)."""
    step = first_step
    smear = {}
    for tile, fused in itertools.product(tile_choices, form_choices):
        batch.set_option(HV_OPT_SMEAR_TILE, tile)
        batch.set_option(HV_OPT_SMEAR_FUSED, fused)
        batch.run_steps(skyrmes, params, step, 1)
        step += 1
        smear[(tile, fused)] = _stage_ms(batch, ("smear",), skyrmes, params, step, steps_per_trial)
        step += steps_per_trial
    tile, fused = min(smear, key=smear.get)
    batch.set_option(HV_OPT_SMEAR_TILE, tile)
    batch.set_option(HV_OPT_SMEAR_FUSED, fused)
    part = {}
    for sort_every in sort_choices:
        batch.set_option(HV_OPT_SORT_EVERY, sort_every)
        batch.run_steps(skyrmes, params, step, 1)
        step += 1
        n = max(sort_every, steps_per_trial)
        part[sort_every] = _stage_ms(batch, ("particle_pass",), skyrmes, params, step, n)
        step += n
    sort_every = min(part, key=part.get)
    batch.set_option(HV_OPT_SORT_EVERY, sort_every)
    timings = {(s, t, f): part[s] + smear[(t, f)] for s in sort_choices for (t, f) in smear}
    return {"sort_every": sort_every, "smear_tile": tile, "smear_fused": fused}, timings, step


def retune_batch_sort(batch, skyrmes, params, first_step: int, steps_per_trial: int = 8,
                      sort_choices=(0, 4, 8, 16)):
    """The sort cadence alone, again, on a running ensemble (the particles'
    order decays along a run, so the cadence that is fastest at its start is
    not the one for later steps): each candidate runs one untimed step, then
    max(cadence, steps_per_trial) steps compared on the particle stage with
    its sorts.  Returns (best sort_every, {cadence: ms/step}, next
    first_step)."""
    step = first_step
    part = {}
    for sort_every in sort_choices:
        batch.set_option(HV_OPT_SORT_EVERY, sort_every)
        batch.run_steps(skyrmes, params, step, 1)
        step += 1
        n = max(sort_every, steps_per_trial)
        part[sort_every] = _stage_ms(batch, ("particle_pass",), skyrmes, params, step, n)
        step += n
    best = min(part, key=part.get)
    batch.set_option(HV_OPT_SORT_EVERY, best)
    return best, part, step


def retune_batch_steps(steps_per_trial: int = 8, sort_choices=(0, 4, 8, 16)) -> int:
    """Steps one retune_batch_sort round runs."""
    return sum(1 + max(c, steps_per_trial) for c in sort_choices)


SINGLE_SORT_CHOICES = (0, 2, 4, 8, 16, 32)


class _Steps:
    """Runs steps of one context, keeping each step's stats and SOR
    iterations (the tuning trials are steps of the run itself)."""

    def __init__(self, sim, skyrme, params):
        self.sim, self.skyrme, self.params = sim, skyrme, params
        self.stats, self.its = [], []

    def run(self, step, n, stages=None):
        if stages:
            self.sim.profile_enable(True)
        s, it = self.sim.run_steps(self.skyrme, self.params, step, n)
        self.stats.append(s)
        self.its.append(it)
        if not stages:
            return None
        prof = self.sim.profile_read()
        self.sim.profile_enable(False)
        return sum(prof[k][0] for k in stages) / n


def tune_single(sim, skyrme, params, first_step: int, steps_per_trial: int = 16,
                sort_choices=SINGLE_SORT_CHOICES, tile_choices=(), form_choices=(), _runner=None):
    """The same search for one context (hvamd.Simulation), on the live run:
    each candidate sort cadence advances the run by max(cadence,
    steps_per_trial) steps after one untimed step and is compared on the
    particle pass's device time per step (the pass with its re-sorts, which
    is all a cadence changes); tile_choices / form_choices, when given, are
    compared on the smearing's time.  Returns (best {sort_every[, smear_tile,
    smear_fused]}, timings {knob: {value: ms/step}}, next first_step)."""
    r = _runner or _Steps(sim, skyrme, params)
    step = first_step
    best, timings = {}, {}
    if tile_choices or form_choices:
        smear = {}
        for tile, fused in itertools.product(tile_choices or (64,), form_choices or (0,)):
            sim.set_option(HV_OPT_SMEAR_TILE, tile)
            sim.set_option(HV_OPT_SMEAR_FUSED, fused)
            r.run(step, 1)
            step += 1
            smear[(tile, fused)] = r.run(step, steps_per_trial, ("smear",))
            step += steps_per_trial
        tile, fused = min(smear, key=smear.get)
        sim.set_option(HV_OPT_SMEAR_TILE, tile)
        sim.set_option(HV_OPT_SMEAR_FUSED, fused)
        best.update(smear_tile=tile, smear_fused=fused)
        timings["smear"] = {f"tile{t}_fused{f}": v for (t, f), v in smear.items()}
    part = {}
    for sort_every in sort_choices:
        sim.set_option(HV_OPT_SORT_EVERY, sort_every)
        r.run(step, 1)
        step += 1
        n = max(sort_every, steps_per_trial)
        part[sort_every] = r.run(step, n, ("particle_pass",))
        step += n
    sort_every = min(part, key=part.get)
    sim.set_option(HV_OPT_SORT_EVERY, sort_every)
    best["sort_every"] = sort_every
    timings["particle_pass"] = part
    return best, timings, step


def tuning_steps(steps_per_trial: int = 16, sort_choices=SINGLE_SORT_CHOICES) -> int:
    """Steps one tune_single round of the sort cadence runs."""
    return sum(1 + max(c, steps_per_trial) for c in sort_choices)


def run_retuned(sim, skyrme, params, first_step: int, nsteps: int, retune_every: int = 500, **tune_args):
    """nsteps steps of the loop from first_step, with the sort cadence tuned
    (tune_single) at the start and again after every retune_every steps; the
    tuning trials are steps of the run.  Returns (stats[nsteps, 4],
    sor_iters[nsteps], [(step, best, timings)] of each tuning)."""
    import numpy as np
    r = _Steps(sim, skyrme, params)
    step, end = first_step, first_step + nsteps
    need = tuning_steps(tune_args.get("steps_per_trial", 16), tune_args.get("sort_choices", SINGLE_SORT_CHOICES))
    log = []
    while step < end:
        if end - step > need:
            t0 = step
            best, timings, step = tune_single(sim, skyrme, params, step, _runner=r, **tune_args)
            log.append((t0, best, timings))
        n = min(retune_every, end - step)
        if n > 0:
            r.run(step, n)
            step += n
    return np.concatenate(r.stats), np.concatenate(r.its), log
