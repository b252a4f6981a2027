"""Bit-exact parity of the HIP path (through the C ABI) with the oracle.

Every comparison is exact: integer outputs, int16 samples, and the fp32 bits
of the RD cost.  Sizes are small enough for the oracle to finish in seconds.
"""
import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture

pytestmark = pytest.mark.gpu

RECORD_FIELDS = [n for n in evxa.RECORD_DTYPE.names if not n.startswith("reserved")]


@pytest.fixture(autouse=True, params=[evxa.MODE_WAVEFRONT, evxa.MODE_SWEEP],
                ids=["wavefront", "sweep"])
def mode(request, ctx):
    """Every test of this module runs in both ways of resolving the MV
    predictor's dependencies; both must give the oracle's outputs."""
    before = ctx.tuning()[0]
    ctx.set_tuning(mode=request.param)
    yield request.param
    ctx.set_tuning(**before)


def _wide_picture(rng, w, h, lo, hi):
    return (rng.integers(lo, hi, (h, w)).astype(np.int16),
            rng.integers(lo, hi, (h // 2, w // 2)).astype(np.int16),
            rng.integers(lo, hi, (h // 2, w // 2)).astype(np.int16))


def _upload(stream, slot, pic):
    stream.upload(slot, *pic)


def assert_records_equal(got, want, fields=RECORD_FIELDS):
    for f in fields:
        bad = np.nonzero(got[f] != want[f])[0]
        assert bad.size == 0, (f"field {f}: {bad.size} MBs differ, first {bad[:8]} "
                               f"got {got[f][bad[:4]]} want {want[f][bad[:4]]}")


@pytest.mark.parametrize("lo,hi", [(0, 256), (-3000, 3000)])
def test_interp_blocks(ctx, oracle, lo, hi):
    rng = np.random.default_rng(11 + abs(lo))
    w, h = 96, 64
    pic = _wide_picture(rng, w, h, lo, hi)
    s = ctx.open_stream(w, h, 1)
    _upload(s, 0, pic)
    P = Picture(*pic)
    xs, ys, dxs, dys, qs = [], [], [], [], []
    for bx in (0, 1, 7, 40, w - 17, w - 16):
        for by in (0, 3, 24, h - 16):
            for k, (dx, dy) in enumerate(synth.DIRS):
                for q in (0, 1):
                    xs.append(bx); ys.append(by); dxs.append(dx); dys.append(dy); qs.append(q)
    got = s.interpolate_subpixel_6tap(0, xs, ys, dxs, dys, qs)
    for k in range(len(xs)):
        want = oracle.interp(P, xs[k], ys[k], dxs[k], dys[k], qs[k])
        assert np.array_equal(got[k], want), (xs[k], ys[k], dxs[k], dys[k], qs[k])
    s.close()


@pytest.mark.parametrize("amp", [255, 600, 4095])
def test_dct_quant_blocks(ctx, oracle, amp):
    rng = np.random.default_rng(amp)
    n = 64
    src = rng.integers(-amp, amp + 1, (n, 384)).astype(np.int16)
    pred = rng.integers(-amp, amp + 1, (n, 384)).astype(np.int16)
    src[0] = pred[0]                       # zero residual: zero out (test_transform.cpp:169-196)
    quality = rng.integers(0, 64, n).astype(np.int32)
    quality[:8] = [0, 1, 8, 12, 15, 16, 31, 63]
    co, lv, qp, var = ctx.dct_quant_blocks(src, pred, quality)
    for k in range(n):
        c = oracle.sub_transform(src[k], pred[k])
        assert np.array_equal(co[k], c), k
        assert var[k] == oracle.variance2(c), k
        q = oracle.block_qp(int(quality[k]), c)
        assert qp[k] == q, k
        assert np.array_equal(lv[k], oracle.quantize(q, c)), k
    assert not co[0].any() and not lv[0].any()


def _pair(seed, w, h, noise=True):
    ref, cur = synth.pair_c0(seed, w, h)
    if not noise:
        rng = np.random.default_rng(seed)
        hb, wb = h // 16, w // 16
        mvx = rng.integers(-16, 17, (hb, wb))
        mvy = rng.integers(-16, 17, (hb, wb))
        sp = rng.integers(-1, 16, (hb, wb))
        cur = synth.move_blocks(rng, ref, mvx, mvy, sp, noise_sigma=0.0)
    return ref, cur


@pytest.mark.parametrize("quality", [8, 9, 12, 14])
def test_search_blocks(ctx, oracle, quality):
    w, h = 192, 128
    ref, cur = _pair(quality, w, h)
    s = ctx.open_stream(w, h, 2)
    _upload(s, 0, cur)
    _upload(s, 1, ref)
    R, S = Picture(*ref), Picture(*cur)
    rng = np.random.default_rng(quality)
    xs, ys, pxs, pys = [], [], [], []
    for by in range(h // 16):
        for bx in range(w // 16):
            for pm in ((0, 0), tuple(rng.integers(-20, 21, 2)), (200, -300)):
                xs.append(bx * 16); ys.append(by * 16); pxs.append(pm[0]); pys.append(pm[1])
    got = s.perform_inter_search(0, 1, quality, xs, ys, pxs, pys)
    for k in range(len(xs)):
        r = oracle.inter_search(S.block(xs[k], ys[k]), R, xs[k], ys[k], quality, pxs[k], pys[k])
        for f in evxa.SEARCH_DTYPE.names:
            assert got[f][k] == getattr(r, f), (k, f, got[k], xs[k], ys[k], pxs[k], pys[k])
    s.close()


def _stretch(pics, lo, hi):
    """Map the samples of all `pics` affinely onto [lo, hi] (same map for all)."""
    a = min(int(c.min()) for p in pics for c in p)
    b = max(int(c.max()) for p in pics for c in p)
    return [tuple(((c.astype(np.int64) - a) * (hi - lo) // (b - a) + lo).astype(np.int16)
                  for c in p) for p in pics]


# Sample ranges that select the metric paths.  A source / reference pair
# takes the packed 16-bit path when their samples together span at most 2048
# values (max - min <= 2047) within [-8192, 8191], else the 32-bit path:
#   packed: 8-bit content as generated; [0, 2047] and [1, 2048] (span 2047,
#     the largest differences); [-3000, -1000]; and the limits of the window,
#     [6144, 8191] and [-8192, -6145], where half-pel samples come closest
#     to the int16 range;
#   32-bit: [0, 2048] (span 2048, one value too many), [6145, 8192] (one past
#     the top), and [-32000, 32000] (differences beyond int16, which the
#     Hadamard input truncates like analysis.h:102).
RANGE_CASES = [None, (0, 2047), (1, 2048), (-3000, -1000), (6144, 8191), (-8192, -6145),
               (0, 2048), (6145, 8192), (-32000, 32000)]


@pytest.mark.parametrize("span", RANGE_CASES)
def test_search_blocks_sample_ranges(ctx, oracle, span):
    w, h = 128, 96
    ref, cur = _pair(77, w, h)
    if span is not None:
        ref, cur = _stretch([ref, cur], *span)
    s = ctx.open_stream(w, h, 2)
    _upload(s, 0, cur)
    _upload(s, 1, ref)
    R, S = Picture(*ref), Picture(*cur)
    rng = np.random.default_rng(5)
    xs, ys, pxs, pys = [], [], [], []
    for by in range(h // 16):
        for bx in range(w // 16):
            pm = tuple(rng.integers(-12, 13, 2)) if (bx + by) % 2 else (0, 0)
            xs.append(bx * 16); ys.append(by * 16); pxs.append(pm[0]); pys.append(pm[1])
    got = s.perform_inter_search(0, 1, 8, xs, ys, pxs, pys)
    for k in range(len(xs)):
        r = oracle.inter_search(S.block(xs[k], ys[k]), R, xs[k], ys[k], 8, pxs[k], pys[k])
        for f in evxa.SEARCH_DTYPE.names:
            assert got[f][k] == getattr(r, f), (k, f, got[k], xs[k], ys[k], pxs[k], pys[k])
    s.close()


def test_search_blocks_disjoint_ranges(ctx, oracle):
    """Each picture alone spans 2048 values, the pair together does not: the
    pair must take the 32-bit path."""
    w, h = 128, 96
    ref, cur = _pair(78, w, h)
    ref = _stretch([ref], 3000, 5047)[0]
    cur = _stretch([cur], 0, 2047)[0]
    s = ctx.open_stream(w, h, 2)
    _upload(s, 0, cur)
    _upload(s, 1, ref)
    R, S = Picture(*ref), Picture(*cur)
    xs = [bx * 16 for by in range(h // 16) for bx in range(w // 16)]
    ys = [by * 16 for by in range(h // 16) for bx in range(w // 16)]
    pxs = [3 * (k % 5) - 6 for k in range(len(xs))]
    pys = [2 * (k % 3) - 2 for k in range(len(xs))]
    got = s.perform_inter_search(0, 1, 8, xs, ys, pxs, pys)
    for k in range(len(xs)):
        r = oracle.inter_search(S.block(xs[k], ys[k]), R, xs[k], ys[k], 8, pxs[k], pys[k])
        for f in evxa.SEARCH_DTYPE.names:
            assert got[f][k] == getattr(r, f), (k, f)
    s.close()


@pytest.mark.parametrize("span", RANGE_CASES[1:])
def test_frame_p_sample_ranges(ctx, oracle, span):
    frames = _stretch(synth.sequence_c1(91, 3, 128, 96, local_range=10), *span)
    cur, refs = frames[2], [frames[1], frames[0]]
    rec, co, lv = _run_p(ctx, cur, refs, 10)
    orec, oco, olv = oracle.analyze_frame(1, 10, Picture(*cur), [Picture(*r) for r in refs])
    assert_records_equal(rec, orec)
    assert np.array_equal(co, oco)
    assert np.array_equal(lv, olv)


def _run_p(ctx, cur, refs, quality, prev_frames=()):
    h, w = cur[0].shape
    s = ctx.open_stream(w, h, 1 + len(refs))
    _upload(s, 0, cur)
    for k, r in enumerate(refs):
        _upload(s, 1 + k, r)
    s.analyze_p(quality, 0, list(range(1, 1 + len(refs))))
    out = s.results()
    s.close()
    return out


@pytest.mark.parametrize("quality,noise", [(8, True), (12, True), (8, False), (13, False),
                                           (0, True), (63, True), (63, False)])
def test_frame_p_single_ref(ctx, oracle, quality, noise):
    w, h = 160, 96
    ref, cur = _pair(100 + quality, w, h, noise)
    rec, co, lv = _run_p(ctx, cur, [ref], quality)
    orec, oco, olv = oracle.analyze_frame(1, quality, Picture(*cur), [Picture(*ref)])
    assert_records_equal(rec, orec)
    assert np.array_equal(co, oco)
    assert np.array_equal(lv, olv)
    if not noise:
        assert (rec["block_type"] & 4).any(), "expected copy blocks on noise-free content"


@pytest.mark.parametrize("n_refs,quality", [(2, 8), (4, 8), (3, 31), (4, 0)])
def test_frame_p_multi_ref(ctx, oracle, n_refs, quality):
    frames = synth.sequence_c1(7, n_refs + 1, 160, 96, local_range=12)
    cur = frames[-1]
    refs = [frames[-2 - k] for k in range(n_refs)]
    rec, co, lv = _run_p(ctx, cur, refs, quality)
    orec, oco, olv = oracle.analyze_frame(1, quality, Picture(*cur), [Picture(*r) for r in refs])
    assert_records_equal(rec, orec)
    assert np.array_equal(co, oco)
    assert np.array_equal(lv, olv)


def test_frame_p_temporal_chain(ctx, oracle):
    """Three P frames; frames 2 and 3 use the previous frame's table (H14)."""
    frames = synth.sequence_c1(21, 4, 128, 96, local_range=10)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 4)
    for k, f in enumerate(frames):
        _upload(s, k, f)
    prev = None
    for t in range(1, 4):
        refs = list(range(t - 1, max(-1, t - 3), -1))
        s.analyze_p(9, t, refs)
        rec, co, lv = s.results()
        orec, oco, olv = oracle.analyze_frame(1, 9, Picture(*frames[t]),
                                              [Picture(*frames[r]) for r in refs], prev=prev)
        assert_records_equal(rec, orec)
        assert np.array_equal(co, oco) and np.array_equal(lv, olv)
        prev = orec
    s.close()


@pytest.mark.parametrize("quality", [9, 13, 0, 40, 63])
def test_frame_b(ctx, oracle, quality):
    frames = synth.sequence_c1(31, 3, 160, 96, local_range=12)
    fwd, cur, bwd = frames
    h, w = cur[0].shape
    s = ctx.open_stream(w, h, 3)
    for k, f in enumerate((cur, fwd, bwd)):
        _upload(s, k, f)
    s.analyze_b(quality, 0, 1, 2, target_ids=[0, 1])
    rec, co, lv = s.results()
    orec, oco, olv = oracle.analyze_frame(2, quality, Picture(*cur), [Picture(*fwd), Picture(*bwd)])
    assert_records_equal(rec, orec)
    assert np.array_equal(co, oco) and np.array_equal(lv, olv)
    s.close()


def test_gop8_hierarchical_sequence(ctx, oracle):
    """A GOP-8 sequence with a 3-level B pyramid in decode order: P anchors
    chain their temporal tables, Bs use forward/backward/bidirectional
    prediction at the pyramid's qualities."""
    from chicago_amd import gop
    n = 10
    frames = synth.sequence_c2(61, n, 128, 96, local_range=12)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, n)
    for k, f in enumerate(frames):
        _upload(s, k, f)
    ptables = {}
    for r in gop.gop8_schedule(n, 12):
        if r.kind == "I":
            continue
        cur = Picture(*frames[r.index])
        if r.kind == "P":
            if not r.temporal_mvp:
                s.reset_temporal()
            s.analyze_p(r.quality, r.index, r.refs)
            prev = ptables.get(r.refs[0]) if r.temporal_mvp else None
            want = oracle.analyze_frame(1, r.quality, cur, [Picture(*frames[r.refs[0]])], prev=prev)
        else:
            s.analyze_b(r.quality, r.index, r.refs[0], r.refs[1], target_ids=[0, 1])
            want = oracle.analyze_frame(2, r.quality, cur, [Picture(*frames[k]) for k in r.refs],
                                        target_ids=[0, 1])
        rec, co, lv = s.results()
        assert_records_equal(rec, want[0])
        assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), r
        if r.kind == "P":
            ptables[r.index] = want[0]
    s.close()


def test_reference_gop4_sequence(ctx, oracle):
    """The reference encoder's own GOP (P anchor + 3 Bs, reference B in the
    middle, evx3enc.cpp:279-383) over 11 frames, flush included, in decode
    order: records carry the picture-buffer slots as prediction targets, P
    anchors chain the previous non-B frame's table."""
    from chicago_amd import gop
    n = 11
    frames = synth.sequence_c2(62, n, 112, 80, local_range=12)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, n)
    for k, f in enumerate(frames):
        _upload(s, k, f)
    tables = {}
    kinds = set()
    for r in gop.reference_schedule(n, 10):
        if r.kind == "I":
            continue
        cur = Picture(*frames[r.index])
        refs = [Picture(*frames[k]) for k in r.refs]
        if r.kind == "P":
            prev = tables.get(r.temporal_from)
            if prev is None:
                s.reset_temporal()
            s.analyze_p(r.quality, r.index, r.refs, target_ids=r.ref_slots)
            want = oracle.analyze_frame(1, r.quality, cur, refs, prev=prev, target_ids=r.ref_slots)
            tables[r.index] = want[0]
        else:
            s.analyze_b(r.quality, r.index, r.refs[0], r.refs[1], target_ids=r.ref_slots)
            want = oracle.analyze_frame(2, r.quality, cur, refs, target_ids=r.ref_slots)
        rec, co, lv = s.results()
        assert_records_equal(rec, want[0])
        assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), r
        kinds.add((r.kind, r.is_reference, r.quality))
    assert kinds == {("P", True, 10), ("B", True, 11), ("B", False, 12)}
    s.close()


def test_batch_uneven_streams_match_alone(ctx, oracle):
    """Streams of different sizes in one fused launch give the same results
    as each analysed alone, and as the oracle."""
    sizes = [(64, 48), (160, 96), (96, 64), (128, 128)]
    streams, descs, expect = [], [], []
    for k, (w, h) in enumerate(sizes):
        frames = synth.sequence_c1(40 + k, 3, w, h, local_range=8)
        s = ctx.open_stream(w, h, 3)
        for i, f in enumerate(frames):
            _upload(s, i, f)
        streams.append(s)
        nref = 1 + k % 2
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


def test_graph_replays_match_eager(ctx, oracle):
    frames = synth.sequence_c1(55, 2, 128, 64, local_range=8)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 2)
    for i, f in enumerate(frames):
        _upload(s, i, f)
    d = evxa.frame_desc(evxa.FRAME_P, 10, 1, [0], temporal_mvp=False, update_temporal=False)
    s.analyze(d)
    eager = s.results()
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
