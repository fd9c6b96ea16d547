"""The ensemble autotuner's search over a stand-in batch whose stage times
are known functions of the knobs, with a SOR stage that grows along the run
(which must not sway the choice).  CPU only."""
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "hyper-vlasov_amd"))

from hvamd._lib import HV_OPT_SMEAR_FUSED, HV_OPT_SMEAR_TILE, HV_OPT_SORT_EVERY  # noqa: E402
from hvamd.autotune import tune_batch  # noqa: E402


class StandInBatch:
    """Smearing is fastest at tile 32, two-pass; the particle stage is
    fastest with sorting every 8 steps (a sort costs 5, an unsorted pass 3, a
    sorted one 1); the SOR stage grows with the step number."""

    def __init__(self):
        self.opt = {}
        self.steps = []
        self.prof = None
        self.since_sort = 1 << 30

    def set_option(self, option, value):
        self.opt[option] = value
        if option == HV_OPT_SORT_EVERY:
            self.since_sort = 1 << 30

    def profile_enable(self, on):
        self.prof = {"smear": [0.0, 0], "particle_pass": [0.0, 0], "sor_stage": [0.0, 0]} if on else None

    def profile_read(self):
        return {k: tuple(v) for k, v in self.prof.items()}

    def run_steps(self, skyrmes, params, first, n):
        self.steps.append((first, n))
        for s in range(first, first + n):
            sort_every = self.opt.get(HV_OPT_SORT_EVERY, 0)
            part = 3.0
            if sort_every > 0:
                part = 1.0
                if self.since_sort >= sort_every:
                    part += 5.0
                    self.since_sort = 0
                self.since_sort += 1
            smear = abs(self.opt[HV_OPT_SMEAR_TILE] - 32) / 8 + 2 * self.opt[HV_OPT_SMEAR_FUSED] + 1
            if self.prof is not None:
                for k, v in (("smear", smear), ("particle_pass", part), ("sor_stage", 100.0 + 10 * s)):
                    self.prof[k][0] += v
                    self.prof[k][1] += 1


def test_tune_batch_picks_fastest_stages_and_leaves_them_set():
    b = StandInBatch()
    best, timings, nxt = tune_batch(b, [], [], 5, steps_per_trial=2, sort_choices=(0, 8), tile_choices=(16, 32),
                                    form_choices=(0, 1))
    assert best == {"sort_every": 8, "smear_tile": 32, "smear_fused": 0}
    assert b.opt == {HV_OPT_SORT_EVERY: 8, HV_OPT_SMEAR_TILE: 32, HV_OPT_SMEAR_FUSED: 0}
    assert set(timings) == {(s, t, f) for s in (0, 8) for t in (16, 32) for f in (0, 1)}
    # smear trials 4 x (1 + 2) steps, then sort trials (1 + 2) and (1 + 8)
    assert nxt == 5 + 4 * 3 + 3 + 9
    firsts = [f for f, _ in b.steps]
    assert firsts == sorted(firsts) and firsts[0] == 5
    # the sort-8 trial's cost includes one sort over its 8 steps
    assert abs(timings[(8, 32, 0)] - (1.0 + 5.0 / 8 + 1.0)) < 1e-12


class StandInSim(StandInBatch):
    """One context: run_steps returns per-step stats and SOR iterations (the
    step number, so the test can check every step was run once, in order)."""

    def run_steps(self, skyrme, params, first, n):
        import numpy as np
        super().run_steps([], [], first, n)
        steps = np.arange(first, first + n)
        return np.repeat(steps[:, None], 4, axis=1).astype(float), steps.astype(np.int32)


def test_tune_single_picks_the_cheapest_cadence_on_the_particle_pass():
    from hvamd.autotune import tune_single
    s = StandInSim()
    s.set_option(HV_OPT_SMEAR_TILE, 64)
    s.set_option(HV_OPT_SMEAR_FUSED, 0)
    best, timings, nxt = tune_single(s, None, None, 10, steps_per_trial=4, sort_choices=(0, 2, 8, 16))
    # sorting every 16 steps: 1 + 5/16 per step, the cheapest of 3, 1 + 5/4, 1 + 5/8, 1 + 5/16
    assert best == {"sort_every": 16}
    assert s.opt[HV_OPT_SORT_EVERY] == 16
    assert abs(timings["particle_pass"][8] - (1.0 + 5.0 / 8)) < 1e-12
    assert nxt == 10 + (1 + 4) + (1 + 4) + (1 + 8) + (1 + 16)


def test_run_retuned_runs_every_step_once_in_order_and_retunes():
    from hvamd.autotune import run_retuned, tuning_steps
    s = StandInSim()
    s.set_option(HV_OPT_SMEAR_TILE, 64)
    s.set_option(HV_OPT_SMEAR_FUSED, 0)
    kw = dict(steps_per_trial=4, sort_choices=(0, 8))
    need = tuning_steps(4, (0, 8))
    stats, its, log = run_retuned(s, None, None, 3, 200, retune_every=50, **kw)
    assert list(its) == list(range(3, 203))
    assert stats.shape == (200, 4)
    assert [t for t, _, _ in log][:2] == [3, 3 + need + 50]
    assert all(b["sort_every"] == 8 for _, b, _ in log)
