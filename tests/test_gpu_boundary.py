"""Admission by HBM budget and per-stream completion on a real context,
through the C ABI."""
import time

import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture

from test_gpu_parity import assert_records_equal

pytestmark = pytest.mark.gpu


def test_hbm_budget_refuses_stream_open():
    """A context whose budget holds two streams of this size refuses a third
    with EVXA_ERROR_CAPACITY_LIMIT (7) and charges nothing for it; closing
    one admits it again."""
    w, h, slots = 128, 96, 3
    one = evxa.stream_bytes(w, h, slots)
    ctx = evxa.Context(0, hbm_budget=2 * one + one // 2)
    try:
        a = ctx.open_stream(w, h, slots)
        b = ctx.open_stream(w, h, slots)
        used = ctx.bytes_used
        assert used == 2 * one
        with pytest.raises(evxa.EvxaError) as e:
            ctx.open_stream(w, h, slots)
        assert e.value.code == 7
        assert ctx.bytes_used == used
        # A smaller stream still fits the remaining half.
        small = ctx.open_stream(16, 16, 1)
        assert ctx.bytes_used == used + evxa.stream_bytes(16, 16, 1)
        small.close()
        a.close()
        assert ctx.bytes_used == one
        c = ctx.open_stream(w, h, slots)
        assert ctx.bytes_used == 2 * one
        c.close()
        b.close()
        assert ctx.bytes_used == 0
    finally:
        ctx.close()


def _open(ctx, frames):
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, len(frames))
    for k, f in enumerate(frames):
        s.upload(k, *f)
    return s


@pytest.mark.parametrize("mode", [evxa.MODE_WAVEFRONT, evxa.MODE_SWEEP])
def test_stream_wait_covers_each_kind_of_job(ctx, oracle, mode):
    """evxa_stream_wait returns once the stream's own jobs are done: after
    a frame, a batch, a reconstruction and a graph replay the stream reports
    done and its outputs equal the oracle's."""
    before = ctx.tuning()[0]
    ctx.set_tuning(mode=mode)
    try:
        fa = synth.sequence_c1(81, 3, 160, 96, local_range=10)
        fb = synth.sequence_c1(82, 3, 96, 64, local_range=10)
        a, b = _open(ctx, fa), _open(ctx, fb)
        want_a = oracle.analyze_frame(1, 9, Picture(*fa[2]), [Picture(*fa[1])])
        want_b = oracle.analyze_frame(1, 9, Picture(*fb[2]), [Picture(*fb[1]), Picture(*fb[0])])
        da = evxa.frame_desc(evxa.FRAME_P, 9, 2, [1], temporal_mvp=False, update_temporal=False)
        db = evxa.frame_desc(evxa.FRAME_P, 9, 2, [1, 0], temporal_mvp=False,
                             update_temporal=False)
        assert b.query()                     # nothing queued yet
        a.analyze(da)
        b.analyze(db)
        b.wait()
        assert b.query()
        a.wait()
        assert a.query()
        assert_records_equal(a.results()[0], want_a[0])
        assert_records_equal(b.results()[0], want_b[0])
        ctx.analyze_batch([a, b], [da, db])
        a.wait()
        b.wait()
        assert a.query() and b.query()
        assert_records_equal(b.results()[0], want_b[0])
        a.reconstruct(0)
        a.wait()
        assert a.query()
        gid = ctx.graph_create([b], [db])
        ctx.graph_launch(gid)
        b.wait()
        assert b.query()
        rec, co, lv = b.results()
        assert_records_equal(rec, want_b[0])
        assert np.array_equal(co, want_b[1]) and np.array_equal(lv, want_b[2])
        ctx.graph_destroy(gid)
        a.close()
        b.close()
    finally:
        ctx.set_tuning(**before)


def test_stream_wait_after_server_launch(ctx, oracle):
    """A frame submitted to the server: the stream's wait returns once its
    launch is done, and the completion then arrives."""
    f = synth.sequence_c1(83, 2, 128, 64, local_range=8)
    s = _open(ctx, f)
    want = oracle.analyze_frame(1, 10, Picture(*f[1]), [Picture(*f[0])])
    d = evxa.frame_desc(evxa.FRAME_P, 10, 1, [0], temporal_mvp=False, update_temporal=False)
    with evxa.Server(ctx, max_inflight=2) as srv:
        srv.submit(s, d, tag=5)
        t0 = time.time()
        done = []
        while len(done) == 0 and time.time() - t0 < 10:
            s.wait()
            done = srv.poll()
        assert [int(c["tag"]) for c in done] == [5] and int(done[0]["status"]) == 0
    assert s.query()
    assert_records_equal(s.results()[0], want[0])
    s.close()
