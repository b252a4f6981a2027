"""Sweep mode (fixpoint of the MV predictor) on the GPU, through the C ABI.

The records, coefficients and levels must equal the oracle's sequential
frame driver bit for bit, and the sweeps themselves are deterministic: the
MBs each sweep listed and searched must equal the CPU model of the same
iteration (oracle.sweep_model, over the oracle's searches).
"""
import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture, sweep_model

from test_gpu_parity import assert_records_equal

pytestmark = pytest.mark.gpu


@pytest.fixture
def sweep_ctx(ctx):
    before = ctx.tuning()[0]
    ctx.set_tuning(mode=evxa.MODE_SWEEP)
    yield ctx
    ctx.set_tuning(**before)


def _open(ctx, frames):
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, len(frames))
    for k, f in enumerate(frames):
        s.upload(k, *f)
    return s


def _check_frame(ctx, oracle, s, ftype, quality, cur, refs, prev, label, stats=True):
    rec, co, lv = s.results()
    want = oracle.analyze_frame(ftype, quality, Picture(*cur), [Picture(*r) for r in refs],
                                prev=prev, target_ids=None if ftype == 1 else [0, 1])
    assert_records_equal(rec, want[0])
    assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), label
    if stats:
        words, listed, searched = sweep_model(oracle, ftype, quality, Picture(*cur),
                                              [Picture(*r) for r in refs], prev)
        got_l, got_s = ctx.sweep_stats()
        assert (got_l, got_s) == (listed, searched), (label, got_l, got_s, listed, searched)
    return want[0]


# Sweep 0's neighbour reads: the guess table only (deterministic sweeps,
# checked against the model), or also MVs decided earlier in sweep 0, with
# and without waiting for them (timing-dependent sweeps, same results).
READS = [pytest.param((evxa.SWEEP_READS_GUESS, 0), id="guess"),
         pytest.param((evxa.SWEEP_READS_FRESH, 0), id="fresh"),
         pytest.param((evxa.SWEEP_READS_FRESH, 30), id="fresh-wait30us")]


@pytest.mark.parametrize("reads", READS)
@pytest.mark.parametrize("waves", [1, 2])
def test_sweep_p_temporal_chain(sweep_ctx, oracle, waves, reads):
    """P frames with 1-3 references; from the second frame on the temporal
    table is both a predictor candidate and sweep 0's guess."""
    ctx = sweep_ctx
    ctx.set_tuning(waves_per_search=waves, mode=evxa.MODE_SWEEP, sweep_reads=reads[0],
                   sweep_wait_us=reads[1])
    frames = synth.sequence_c1(17 + waves, 5, 192, 128, local_range=12)
    s = _open(ctx, frames)
    prev = None
    for t in range(1, 5):
        refs = list(range(t - 1, max(-1, t - 4), -1))
        s.analyze_p(8 + t, t, refs)
        prev = _check_frame(ctx, oracle, s, 1, 8 + t, frames[t], [frames[r] for r in refs], prev,
                            f"frame {t}", stats=reads[0] == evxa.SWEEP_READS_GUESS)
    s.close()


@pytest.mark.parametrize("reads", READS)
@pytest.mark.parametrize("waves", [1, 2])
def test_sweep_b(sweep_ctx, oracle, waves, reads):
    ctx = sweep_ctx
    ctx.set_tuning(waves_per_search=waves, mode=evxa.MODE_SWEEP, sweep_reads=reads[0],
                   sweep_wait_us=reads[1])
    fwd, cur, bwd = synth.sequence_c1(33, 3, 160, 96, local_range=12)
    s = _open(ctx, [cur, fwd, bwd])
    for q in (9, 13):
        s.analyze_b(q, 0, 1, 2, target_ids=[0, 1])
        _check_frame(ctx, oracle, s, 2, q, cur, [fwd, bwd], None, f"B q{q}",
                     stats=reads[0] == evxa.SWEEP_READS_GUESS)
    s.close()


@pytest.mark.parametrize("quality", [0, 31, 63])
def test_sweep_extreme_qualities(sweep_ctx, oracle, quality):
    """P (two references, then the temporal table) and B frames at the ends
    of the quality range, where the copy threshold and lambda are extreme;
    the sweeps as the model lists them."""
    ctx = sweep_ctx
    ctx.set_tuning(waves_per_search=1, mode=evxa.MODE_SWEEP, sweep_reads=evxa.SWEEP_READS_GUESS,
                   sweep_wait_us=0)
    frames = synth.sequence_c1(40 + quality, 4, 160, 96, local_range=12)
    s = _open(ctx, frames)
    s.analyze_p(quality, 2, [1, 0])
    prev = _check_frame(ctx, oracle, s, 1, quality, frames[2], [frames[1], frames[0]], None, "P")
    s.analyze_p(quality, 3, [2, 1])
    _check_frame(ctx, oracle, s, 1, quality, frames[3], [frames[2], frames[1]], prev, "P temporal")
    s.analyze_b(quality, 1, 0, 2, target_ids=[0, 1])
    _check_frame(ctx, oracle, s, 2, quality, frames[1], [frames[0], frames[2]], None, "B")
    s.close()


def test_sweep_no_motion_single_sweep(sweep_ctx, oracle):
    """A frame equal to its reference: every MB is a copy without motion,
    so sweep 1's check of every MB finds no predictor changed and nothing is
    searched again."""
    ctx = sweep_ctx
    f = synth.sequence_c1(3, 1, 96, 64)[0]
    s = _open(ctx, [f, f])
    s.analyze_p(8, 0, [1])
    rec, _, _ = s.results()
    assert (rec["block_type"] == 4).all()
    assert ctx.sweep_stats() == ([24, 24], [24, 0])
    s.close()


@pytest.mark.parametrize("reads", READS)
def test_sweep_batch_matches_alone(sweep_ctx, oracle, reads):
    """Uneven streams, P and B, in one sweep-mode launch: each equals the
    oracle and the same frame analysed alone."""
    ctx = sweep_ctx
    ctx.set_tuning(mode=evxa.MODE_SWEEP, sweep_reads=reads[0], sweep_wait_us=reads[1])
    sizes = [(64, 48), (160, 96), (96, 64), (128, 128), (208, 112)]
    streams, descs, expect = [], [], []
    for k, (w, h) in enumerate(sizes):
        frames = synth.sequence_c1(60 + k, 3, w, h, local_range=10)
        s = _open(ctx, frames)
        streams.append(s)
        if k == 4:
            descs.append(evxa.frame_desc(evxa.FRAME_B, 12, 1, [0, 2], target_ids=[0, 1]))
            expect.append(oracle.analyze_frame(2, 12, Picture(*frames[1]),
                                               [Picture(*frames[0]), Picture(*frames[2])],
                                               target_ids=[0, 1]))
        else:
            nref = 1 + k % 3
            descs.append(evxa.frame_desc(evxa.FRAME_P, 8 + k, 2, [1, 0][:nref]))
            expect.append(oracle.analyze_frame(1, 8 + k, Picture(*frames[2]),
                                               [Picture(*frames[1]), Picture(*frames[0])][:nref]))
    ctx.analyze_batch(streams, descs)
    for s, (orec, oco, olv) in zip(streams, expect):
        rec, co, lv = s.results()
        assert_records_equal(rec, orec)
        assert np.array_equal(co, oco) and np.array_equal(lv, olv)
    for s, d, (orec, _, _) in zip(streams, descs, expect):
        s.reset_temporal()
        s.analyze(d)
        rec, _, _ = s.results()
        assert_records_equal(rec, orec)
    for s in streams:
        s.close()


def test_sweep_graph_replays_match_eager(sweep_ctx, oracle):
    ctx = sweep_ctx
    frames = synth.sequence_c1(56, 3, 128, 96, local_range=8)
    s = _open(ctx, frames)
    d = evxa.frame_desc(evxa.FRAME_P, 10, 2, [1, 0], temporal_mvp=False, update_temporal=False)
    s.analyze(d)
    eager = s.results()
    want = oracle.analyze_frame(1, 10, Picture(*frames[2]), [Picture(*frames[1]), Picture(*frames[0])])
    assert_records_equal(eager[0], want[0])
    gid = ctx.graph_create([s], [d])
    for _ in range(3):
        ctx.graph_launch(gid)
        ctx.synchronize()
        got = s.results()
        assert_records_equal(got[0], eager[0])
        assert np.array_equal(got[1], eager[1]) and np.array_equal(got[2], eager[2])
        s.analyze(d)              # eager work between replays
        ctx.synchronize()
    ctx.graph_destroy(gid)
    s.close()


def test_sweep_then_wavefront_same_stream(sweep_ctx, oracle):
    """Modes can alternate on one context and stream."""
    ctx = sweep_ctx
    frames = synth.sequence_c1(71, 3, 128, 64, local_range=8)
    s = _open(ctx, frames)
    want = oracle.analyze_frame(1, 9, Picture(*frames[2]), [Picture(*frames[1])])
    for mode in (evxa.MODE_SWEEP, evxa.MODE_WAVEFRONT, evxa.MODE_SWEEP):
        ctx.set_tuning(mode=mode)
        s.reset_temporal()
        s.analyze_p(9, 2, [1])
        rec, co, lv = s.results()
        assert_records_equal(rec, want[0])
        assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2])
    assert ctx.sweep_stats()[0][0] == 32      # sweep 0 lists every MB
    s.close()
