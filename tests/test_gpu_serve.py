"""Continuous batching through the native dispatcher: every frame's results
are identical alone and in the mix, and equal to the oracle's."""
import numpy as np
import pytest

from chicago_amd import evxa, serve, synth
from oracle.oracle import Picture

pytestmark = pytest.mark.gpu

FIELDS = [n for n in evxa.RECORD_DTYPE.names if not n.startswith("reserved")]


def _same(a, b):
    return all(np.array_equal(a[0][f], b[0][f]) for f in FIELDS) and \
        np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def _served_vs_alone(ctx, oracle, sizes, quality, oracle_frames):
    """Each stream's frames 1..4 alone (one launch each), then through the
    dispatcher (every stream's next frame submitted when its previous one is
    read, so launches mix streams and positions): identical, and the first
    stream's first `oracle_frames` frames equal the oracle's chain."""
    n_frames, pool = 5, 4
    content = {sz: synth.sequence_c1(synth.seed_of(9, *sz), pool, *sz, local_range=8)
               for sz in set(sizes)}
    descs = serve.descs_for(pool, quality)
    streams = []
    try:
        for sz in sizes:
            s = ctx.open_stream(synth.align16(sz[0]), synth.align16(sz[1]), pool)
            for k, f in enumerate(content[sz]):
                s.upload(k, *f)
            streams.append(s)
        alone = []
        for s in streams:
            s.reset_temporal()
            out = []
            for k in range(1, n_frames):
                s.analyze(descs[k % pool])
                out.append(s.results())
            alone.append(out)
        prev = None
        f = content[sizes[0]]
        for k in range(1, 1 + oracle_frames):
            want = oracle.analyze_frame(1, quality, Picture(*f[k % pool]),
                                        [Picture(*f[(k - 1) % pool])], prev=prev)
            assert _same(alone[0][k - 1], want), k
            prev = want[0]
        for s in streams:
            s.reset_temporal()
        next_k = [1] * len(streams)
        with evxa.Server(ctx, max_inflight=2, max_frames=3) as srv:
            for i, s in enumerate(streams):
                srv.submit(s, descs[1], i * 100 + 1)
            done = 0
            batches = []
            while done < len(streams) * (n_frames - 1):
                for c in srv.poll():
                    assert c["status"] == 0
                    i, k = divmod(int(c["tag"]), 100)
                    assert k == next_k[i]
                    batches.append(int(c["batch_frames"]))
                    assert c["t_arrival"] <= c["t_launch"] <= c["t_done"]
                    got = streams[i].results()
                    assert _same(got, alone[i][k - 1]), (i, k)
                    done += 1
                    next_k[i] += 1
                    if next_k[i] < n_frames:
                        srv.submit(streams[i], descs[next_k[i] % pool], i * 100 + next_k[i])
            assert max(batches) > 1, "frames were never fused"
    finally:
        for s in streams:
            s.close()


def test_served_frames_match_alone_and_oracle(ctx, oracle):
    # Repeated sizes: the dispatcher fuses frames of one size class.
    _served_vs_alone(ctx, oracle, [(64, 48), (160, 96), (64, 48), (160, 96), (96, 64), (64, 48)],
                     9, 4)


@pytest.mark.parametrize("mode", [pytest.param(evxa.MODE_WAVEFRONT, id="wavefront"),
                                  pytest.param(evxa.MODE_SWEEP, id="sweep")])
def test_served_mix_sizes_match_alone(ctx, oracle, mode):
    """The serving workload's sizes (1280x720, 1920x1080, 2560x1440), in
    each mode."""
    print(f"served mix, mode {mode}", flush=True)
    before = ctx.tuning()[0]
    ctx.set_tuning(mode=mode)
    try:
        _served_vs_alone(ctx, oracle, [(1280, 720), (1920, 1080), (1280, 720), (2560, 1440),
                                       (1920, 1080), (1280, 720)], 8, 1)
    finally:
        ctx.set_tuning(**before)


def test_busy_stream_is_refused_by_direct_calls(ctx):
    """While a server holds a frame of a stream, direct analysis, result reads
    and closing that stream are refused (its outputs belong to the server)."""
    frames = synth.sequence_c1(3, 2, 1280, 720)
    s = ctx.open_stream(1280, 720, 2)
    for k, f in enumerate(frames):
        s.upload(k, *f)
    d = evxa.frame_desc(evxa.FRAME_P, 8, 1, [0])
    with evxa.Server(ctx, max_inflight=1, max_frames=4) as srv:
        srv.submit(s, d, 7)
        srv.submit(s, d, 8)          # queued behind the first frame of its stream
        for call in (lambda: s.analyze(d), s.results, s.close):
            with pytest.raises(evxa.EvxaError) as e:
                call()
            assert e.value.code == 8
        got = []
        while len(got) < 2:
            got += list(srv.poll())
        assert [int(c["tag"]) for c in got] == [7, 8]
        assert all(int(c["batch_frames"]) == 1 for c in got)
    s.results()
    s.close()


def test_serving_loop_small(ctx):
    """A short real-time run of the request loop on small pictures."""
    mix = [(160, 96, 0.5), (128, 64, 0.5)]
    res = serve.run(ctx, seed=3, n_live=6, duration=0.5, mean_life=0.3, pool_frames=4,
                    mix=mix, max_inflight=2, max_frames=8)
    assert res["frames"] > 100
    assert res["latency_ms"]["p50"] > 0
    assert res["admitted_lives"] == res["stream_lives"]
