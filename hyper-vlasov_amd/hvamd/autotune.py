"""Autotuner of an ensemble's smearing tiles and form and its particle-sort
cadence.

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
