"""Autotuner of the analysis launch shape.

The knobs (evxa_set_tuning):
  * the mode: the MB wavefront (each MB waits for its neighbours' final
    MVs) or sweeps to the predictor's fixpoint (no MB waits; MBs whose
    predictor changed are searched again);
  * waves per (MB, reference) search (one wave evaluates a level's 8
    candidates in two passes, two waves in one pass and exchange results
    through LDS);
  * wavefront: work lists per launch (row bands spread over the XCDs' L2
    caches) and the per-list cap of tickets per virtual diagonal that
    staggers the frames of a batch;
  * sweeps: sweep 0's order (raster, or anti-diagonal with a bounded wait
    for the neighbours' MVs, frames staggered by the per-diagonal cap) and
    the tickets claimed at once;
  * both: the reach of the reference window each search stages in LDS
    (a smaller window costs less to stage; levels that leave it load from
    HBM or stage their own).
Which shape is fastest depends on the workload: a single frame is bound by
its MB dependency chain in the wavefront and wants sweeps, a batch of many
frames wants throughput, and content whose predictor changes travel far
(zooms, hierarchical B) needs many sweeps.

`tune` runs a repeatable step of the real workload under each candidate
shape, times its launches with HIP events, requires every shape's outputs to
equal the first shape's bit for bit, and leaves the fastest applied.  The
reference has no counterpart: its analysis runs on CPU threads
(encode.cpp:817-941).
"""
from __future__ import annotations

import itertools

import numpy as np

WAVEFRONT, SWEEP = 1, 2


REACH = 15          # the kernels' largest window reach (the default)


def _wavefront(reach=REACH, **kw):
    return dict(mode=WAVEFRONT, window_reach=reach, **kw)


def _sweep(waves, batch, order=1, wait_us=0, reach=REACH, **kw):
    return dict(mode=SWEEP, waves_per_search=waves, claim_batch=batch, sweep_order=order,
                sweep_reads=2, sweep_wait_us=wait_us, window_reach=reach, **kw)


# Candidate shapes per workload class.  Every candidate sets every knob it
# depends on (set_tuning keeps a knob given as 0).
SINGLE = ([_wavefront(waves_per_search=w, lists=l) for w, l in itertools.product((1, 2), (1, 8))] +
          [_sweep(w, b) for w, b in itertools.product((1, 2), (1, 2))] +
          [_sweep(w, 1, reach=r) for w, r in itertools.product((1, 2), (7, 11))])
BATCH = ([_wavefront(waves_per_search=w, lists=8, cap_pct=c)
          for w, c in itertools.product((1, 2), (100, 150, 250))] +
         [_wavefront(waves_per_search=1, lists=8, cap_pct=150, reach=r) for r in (7, 11)] +
         [_wavefront(waves_per_search=1, lists=8, cap_pct=c, max_groups=2048) for c in (150, 250)] +
         [_sweep(1, 1, order=2, wait_us=w, cap_pct=c) for w, c in
          itertools.product((0, 5, 20), (250, 400))] +
         [_sweep(1, 1, order=2, wait_us=5, cap_pct=250, reach=r) for r in (7, 11)] +
         [_sweep(2, 1, order=2, wait_us=5, cap_pct=250)])
SERVE = ([_wavefront(waves_per_search=w) for w in (1, 2)] + [_sweep(w, 1) for w in (1, 2)] +
         [_sweep(1, 1, reach=r) for r in (7, 11)])


def by_mode(table):
    """Fastest step of each mode in a tune() table."""
    out = {}
    for r in table:
        name = "sweep" if r["tuning"].get("mode") == SWEEP else "wavefront"
        if name not in out or r["step_ms"] < out[name]["step_ms"]:
            out[name] = {"step_ms": r["step_ms"], "tuning": r["tuning"]}
    return out


def _same(a, b):
    if a is None or b is None:
        return True
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def tune(ctx, step, outputs=None, candidates=SINGLE, reps=3, warmup=1, log=None):
    """Time `step()` (which enqueues the launches of one step) under each
    candidate shape; `outputs()` returns arrays to compare across shapes.
    Applies and returns the fastest shape with the per-shape table."""
    table = []
    base = None
    for cand in candidates:
        ctx.set_tuning(**cand)
        for _ in range(warmup):
            step()
        ctx.synchronize()
        ctx.launch_times()
        for _ in range(reps):
            step()
        ctx.synchronize()
        ms = ctx.launch_times()
        step_ms = float(np.sum(ms)) / reps
        out = outputs() if outputs else None
        if base is None:
            base = out
        same = _same(base, out)
        table.append({"tuning": dict(cand), "step_ms": round(step_ms, 4), "bit_identical": same})
        if log:
            log(f"autotune {cand}: {step_ms:.3f} ms/step, bit-identical={same}")
        if not same:
            raise AssertionError(f"launch shape {cand} changed the results")
    best = min(table, key=lambda r: r["step_ms"])
    ctx.set_tuning(**best["tuning"])
    return best, table
