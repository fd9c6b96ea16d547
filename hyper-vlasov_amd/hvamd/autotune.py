"""Autotuner of an ensemble's smearing tiles and form and its particle-sort
cadence.

No knob changes a bit of the results (sorting only reorders particles in
storage; the deposits are exact integer sums; the smearing tile width and form
only regroup the same operations), so they are tuned on the live ensemble:
each candidate advances the run by a few steps and is timed, and the fastest
is left set.  Smearing tiles narrower than the widest that fits the LDS trade
halo re-reads for more blocks; the fused form keeps the D channels in LDS but
walks the planes of a tile in one block, the two-pass form stores them and
runs every (plane, tile) as a block of its own; sorting costs a radix sort per
instance but makes the deposit's LDS boxes compact."""
from __future__ import annotations

import itertools
import time

from ._lib import HV_OPT_SMEAR_FUSED, HV_OPT_SMEAR_TILE, HV_OPT_SORT_EVERY

SORT_CHOICES = (0, 8, 32)
TILE_CHOICES = (16, 32, 64)
FORM_CHOICES = (0, 1)


def tune_batch(batch, skyrmes, params, first_step: int, steps_per_trial: int = 2, sort_choices=SORT_CHOICES,
               tile_choices=TILE_CHOICES, form_choices=FORM_CHOICES):
    """Returns (best {sort_every, smear_tile, smear_fused}, timings
    {(sort, tile, fused): ms/step}, next first_step).  The first trial of each
    setting also pays its one-time costs (tile tables, a first sort), so it is
    run once untimed."""
    timings = {}
    step = first_step
    for sort_every, tile, fused in itertools.product(sort_choices, tile_choices, form_choices):
        batch.set_option(HV_OPT_SORT_EVERY, sort_every)
        batch.set_option(HV_OPT_SMEAR_TILE, tile)
        batch.set_option(HV_OPT_SMEAR_FUSED, fused)
        batch.run_steps(skyrmes, params, step, 1)
        step += 1
        t0 = time.perf_counter()
        batch.run_steps(skyrmes, params, step, steps_per_trial)
        timings[(sort_every, tile, fused)] = 1e3 * (time.perf_counter() - t0) / steps_per_trial
        step += steps_per_trial
    best = min(timings, key=timings.get)
    batch.set_option(HV_OPT_SORT_EVERY, best[0])
    batch.set_option(HV_OPT_SMEAR_TILE, best[1])
    batch.set_option(HV_OPT_SMEAR_FUSED, best[2])
    return {"sort_every": best[0], "smear_tile": best[1], "smear_fused": best[2]}, timings, step
