"""Autotuner host logic with a stand-in context (CPU only)."""
import numpy as np
import pytest

from chicago_amd import autotune


class FakeCtx:
    """Launch times depend on the shape; outputs optionally differ."""

    def __init__(self, times, bad=None):
        self.times, self.bad, self.cur, self.pending = times, bad, None, []

    def set_tuning(self, **kw):
        self.cur = tuple(sorted(kw.items()))

    def synchronize(self):
        pass

    def launch_times(self):
        out, self.pending = np.array(self.pending), []
        return out

    def step(self):
        self.pending.append(self.times[self.cur])

    def outputs(self):
        return (np.array([1, 2, 3]) + (1 if self.cur == self.bad else 0),)


def key(d):
    return tuple(sorted(d.items()))


def test_picks_fastest_and_applies_it():
    cands = [dict(waves_per_search=1), dict(waves_per_search=2)]
    ctx = FakeCtx({key(cands[0]): 5.0, key(cands[1]): 3.0})
    best, table = autotune.tune(ctx, ctx.step, ctx.outputs, cands, reps=3)
    assert best["tuning"] == cands[1] and best["step_ms"] == 3.0
    assert ctx.cur == key(cands[1])
    assert [r["bit_identical"] for r in table] == [True, True]


def test_refuses_a_shape_that_changes_results():
    cands = [dict(waves_per_search=1), dict(waves_per_search=2)]
    ctx = FakeCtx({key(cands[0]): 5.0, key(cands[1]): 3.0}, bad=key(cands[1]))
    with pytest.raises(AssertionError):
        autotune.tune(ctx, ctx.step, ctx.outputs, cands)


def test_candidate_sets_are_valid_shapes():
    for cands in (autotune.SINGLE, autotune.BATCH, autotune.SERVE):
        for c in cands:
            assert c.get("waves_per_search") in (1, 2)
            assert 0 <= c.get("lists", 0) <= 8
