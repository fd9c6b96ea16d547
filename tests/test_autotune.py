"""The ensemble autotuner's search over a stand-in batch whose step time is a
known function of the knobs (CPU only)."""
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "hyper-vlasov_amd"))

from hvamd._lib import HV_OPT_SMEAR_FUSED, HV_OPT_SMEAR_TILE, HV_OPT_SORT_EVERY  # noqa: E402
from hvamd.autotune import tune_batch  # noqa: E402


class StandInBatch:
    """Records options and steps; a step sleeps for a time that is smallest at
    sort 8, tile 32, fused 0."""

    def __init__(self):
        self.opt = {}
        self.steps = []

    def set_option(self, option, value):
        self.opt[option] = value

    def run_steps(self, skyrmes, params, first, n):
        self.steps.append((first, n))
        cost = (abs(self.opt[HV_OPT_SORT_EVERY] - 8) + abs(self.opt[HV_OPT_SMEAR_TILE] - 32) / 8 +
                3 * self.opt[HV_OPT_SMEAR_FUSED])
        time.sleep(n * (0.002 + 0.004 * cost))


def test_tune_batch_picks_fastest_and_leaves_it_set():
    b = StandInBatch()
    best, timings, nxt = tune_batch(b, [], [], 5, steps_per_trial=1, sort_choices=(0, 8), tile_choices=(16, 32),
                                    form_choices=(0, 1))
    assert best == {"sort_every": 8, "smear_tile": 32, "smear_fused": 0}
    assert b.opt == {HV_OPT_SORT_EVERY: 8, HV_OPT_SMEAR_TILE: 32, HV_OPT_SMEAR_FUSED: 0}
    assert set(timings) == {(s, t, f) for s in (0, 8) for t in (16, 32) for f in (0, 1)}
    # each trial: one untimed step then one timed step, consecutive step numbers
    assert nxt == 5 + 8 * 2
    firsts = [f for f, _ in b.steps]
    assert firsts == list(range(5, 5 + 16))
