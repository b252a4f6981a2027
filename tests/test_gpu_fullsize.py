"""Frame analysis at the benchmark workloads' full sizes, in both modes,
bit-exact against the oracle: the 1280x720 single-reference P pair of the
CPU-runnable workload, 1920x1080 P frames with four references and the
temporal table (the plain GPU workload), the first GOP-8 frames of the
3840x2160 zoom sequence with hierarchical B (the GOP workload), and the
batched workload's 64-stream composition in one fused launch."""
import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture, sweep_model

from test_gpu_parity import assert_records_equal

pytestmark = pytest.mark.gpu

MODES = [pytest.param(evxa.MODE_WAVEFRONT, id="wavefront"), pytest.param(evxa.MODE_SWEEP, id="sweep")]


@pytest.fixture
def set_mode(ctx):
    before = ctx.tuning()[0]

    def apply(mode, waves=2, reads=evxa.SWEEP_READS_FRESH):
        ctx.set_tuning(waves_per_search=waves, mode=mode, sweep_reads=reads)
    yield apply
    ctx.set_tuning(**before)


@pytest.mark.parametrize("mode", MODES)
def test_c0_pair_1280x720(ctx, oracle, set_mode, mode):
    # Sweep 0 reads the guess table only, so the sweeps are deterministic
    # and checked against the model below.
    set_mode(mode, reads=evxa.SWEEP_READS_GUESS)
    ref, cur = synth.pair_c0(1, 1280, 720)
    s = ctx.open_stream(1280, 720, 2)
    s.upload(0, *cur)
    s.upload(1, *ref)
    s.analyze_p(8, 0, [1])
    rec, co, lv = s.results()
    s.close()
    want = oracle.analyze_frame(1, 8, Picture(*cur), [Picture(*ref)])
    assert_records_equal(rec, want[0])
    assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2])
    if mode == evxa.MODE_SWEEP:
        _, listed, searched = sweep_model(oracle, 1, 8, Picture(*cur), [Picture(*ref)])
        assert ctx.sweep_stats() == (listed, searched)


@pytest.mark.parametrize("mode", MODES)
def test_c1_1080p_four_refs_temporal(ctx, oracle, set_mode, mode):
    set_mode(mode)
    frames = synth.sequence_c1(1, 7, 1920, 1080)
    h, w = frames[0][0].shape
    assert (w, h) == (1920, 1088)
    s = ctx.open_stream(w, h, len(frames))
    for k, f in enumerate(frames):
        s.upload(k, *f)
    prev = None
    for t in (5, 6):
        refs = [t - 1 - k for k in range(4)]
        if t == 5:
            s.reset_temporal()
        s.analyze_p(8, t, refs)
        rec, co, lv = s.results()
        want = oracle.analyze_frame(1, 8, Picture(*frames[t]), [Picture(*frames[r]) for r in refs],
                                    prev=prev)
        assert_records_equal(rec, want[0])
        assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), t
        prev = want[0]
    s.close()


@pytest.fixture(scope="module")
def c2_case(oracle):
    """The GOP workload's first frames and the oracle's analyses, shared by
    both modes: 3840x2160 zoom sequence, GOP 8 in decode order, the P anchor
    at 8, then the B frames at 4 (references 0 and 8), 2 (0 and 4) and 1
    (0 and 2) at the schedule's qualities."""
    from chicago_amd import gop
    frames = synth.sequence_c2(3, 9)
    roles = [r for r in gop.gop8_schedule(9, 12) if r.index in (8, 4, 2, 1)]
    assert [r.index for r in roles] == [8, 4, 2, 1]
    wants = []
    for r in roles:
        cur = Picture(*frames[r.index])
        refs = [Picture(*frames[k]) for k in r.refs]
        wants.append(oracle.analyze_frame(1, r.quality, cur, refs) if r.kind == "P" else
                     oracle.analyze_frame(2, r.quality, cur, refs, target_ids=[0, 1]))
    return frames, roles, wants


@pytest.mark.parametrize("mode", MODES)
def test_c2_4k_gop8_anchor_and_bs(ctx, c2_case, set_mode, mode):
    frames, roles, wants = c2_case
    set_mode(mode)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, len(frames))
    try:
        for k, f in enumerate(frames):
            s.upload(k, *f)
        for r, want in zip(roles, wants):
            print(f"4K {r.kind} frame {r.index} refs {r.refs} q{r.quality} mode {mode}", flush=True)
            if r.kind == "P":
                s.reset_temporal()
                s.analyze_p(r.quality, r.index, r.refs)
            else:
                s.analyze_b(r.quality, r.index, r.refs[0], r.refs[1], target_ids=[0, 1])
            rec, co, lv = s.results()
            assert_records_equal(rec, want[0])
            assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), r.index
    finally:
        s.close()


@pytest.fixture(scope="module")
def c3_content():
    """One 5-frame sequence per size of the batched workload's mix."""
    return {(w, h): synth.sequence_c1(synth.seed_of(3, w, h), 5, w, h)
            for (w, h) in ((640, 360), (1280, 720), (1920, 1080), (2560, 1440))}


@pytest.mark.parametrize("mode", MODES)
def test_c3_64_streams_batch_matches_alone(ctx, oracle, c3_content, set_mode, mode):
    """The batched workload's composition (16x 640x360, 24x 1280x720, 16x
    1920x1080, 8x 2560x1440; stream i with 1 + i mod 4 references) in one
    fused launch: every stream's outputs equal its launch alone, and the
    largest and smallest streams with four references equal the oracle."""
    set_mode(mode)
    mix = [(640, 360, 16), (1280, 720, 24), (1920, 1080, 16), (2560, 1440, 8)]
    sizes = [(w, h) for w, h, n in mix for _ in range(n)]
    streams, descs = [], []
    try:
        for i, (w, h) in enumerate(sizes):
            frames = c3_content[(w, h)]
            s = ctx.open_stream(synth.align16(w), synth.align16(h), 5)
            for t, f in enumerate(frames):
                s.upload(t, *f)
            streams.append(s)
            descs.append(evxa.frame_desc(evxa.FRAME_P, 8, 4, [3 - k for k in range(1 + i % 4)],
                                         temporal_mvp=False, update_temporal=False))
        print(f"64-stream batch, mode {mode}", flush=True)
        ctx.analyze_batch(streams, descs)
        batch = [s.results() for s in streams]
        for i, (s, d) in enumerate(zip(streams, descs)):
            s.analyze(d)
            alone = s.results()
            assert_records_equal(batch[i][0], alone[0])
            assert np.array_equal(batch[i][1], alone[1]) and np.array_equal(batch[i][2], alone[2]), i
        for i in (3, len(sizes) - 1):       # 640x360 and 2560x1440, four references each
            w, h = sizes[i]
            frames = c3_content[(w, h)]
            want = oracle.analyze_frame(1, 8, Picture(*frames[4]),
                                        [Picture(*frames[3 - k]) for k in range(4)])
            assert_records_equal(batch[i][0], want[0])
            assert np.array_equal(batch[i][1], want[1]) and np.array_equal(batch[i][2], want[2]), i
    finally:
        for s in streams:
            s.close()
