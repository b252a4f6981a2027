"""Reconstruction of analysed frames on the GPU (evxa_reconstruct) against
the oracle's decode-side driver, bit for bit: P frames with 1-4 references,
B frames, levels perturbed as a trellis pass would leave them, samples that
wrap int16, and a closed loop where the reconstruction is the next frame's
reference."""
import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture

pytestmark = pytest.mark.gpu


def _perturb(levels, seed, frac=0.05):
    rng = np.random.default_rng(seed)
    out = levels.copy()
    mask = rng.random(out.shape) < frac
    out[mask] += rng.integers(-2, 3, int(mask.sum())).astype(np.int16)
    return out


def _check_planes(got, want, what):
    for name, g, w in zip("yuv", got, want):
        bad = np.argwhere(g != w)
        assert bad.size == 0, f"{what}: plane {name} differs at {bad[:4].tolist()}"


@pytest.mark.parametrize("n_refs", [1, 2, 4])
def test_reconstruct_p(ctx, oracle, n_refs):
    frames = synth.sequence_c1(40 + n_refs, n_refs + 1, 160, 96, local_range=12)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, n_refs + 2)
    for k, f in enumerate(frames):
        s.upload(k, *f)
    cur = n_refs
    refs = [cur - 1 - k for k in range(n_refs)]
    s.analyze_p(8, cur, refs)
    rec, co, lv = s.results()
    pics = [Picture(*f) for f in frames]
    for tag, levels in (("own", None), ("perturbed", _perturb(lv, n_refs))):
        s.reconstruct(n_refs + 1, levels)
        got = s.download(n_refs + 1)
        want = oracle.reconstruct_frame(1, 8, pics[cur], [pics[r] for r in refs], rec,
                                        lv if levels is None else levels)
        _check_planes(got, want, f"P {n_refs} refs, {tag} levels")
    s.close()


TINY = [(16, 16), (32, 16), (16, 32), (48, 16), (16, 48), (64, 16)]
MODES = [pytest.param(evxa.MODE_WAVEFRONT, id="wavefront"), pytest.param(evxa.MODE_SWEEP, id="sweep")]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("w,h", TINY)
def test_tiny_frames(ctx, oracle, mode, w, h):
    """Frames of one MB row or column (16x16 up to 64x16): P with two
    references, then B, analysed in each mode and reconstructed, against
    the oracle (which matches the reference build at these sizes)."""
    print(f"tiny {w}x{h} mode {mode}", flush=True)
    before = ctx.tuning()[0]
    ctx.set_tuning(mode=mode)
    try:
        frames = synth.sequence_c1(w * 7 + h, 4, w, h, local_range=6)
        pics = [Picture(*f) for f in frames]
        s = ctx.open_stream(w, h, 5)
        try:
            for k, f in enumerate(frames):
                s.upload(k, *f)
            for q in (0, 8, 63):
                for ftype, cur, refs in ((1, 3, [2, 1]), (2, 1, [0, 2])):
                    if ftype == 1:
                        s.analyze_p(q, cur, refs, temporal_mvp=0)
                    else:
                        s.analyze_b(q, cur, refs[0], refs[1], target_ids=[0, 1])
                    rec, co, lv = s.results()
                    want = oracle.analyze_frame(ftype, q, pics[cur], [pics[r] for r in refs],
                                                target_ids=None if ftype == 1 else [0, 1])
                    for f in rec.dtype.names:
                        if not f.startswith("reserved"):
                            assert np.array_equal(rec[f], want[0][f]), (w, h, q, ftype, f)
                    assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), (w, h, q, ftype)
                    s.reconstruct(4)
                    got = s.download(4)
                    planes = oracle.reconstruct_frame(ftype, q, pics[cur], [pics[r] for r in refs],
                                                      rec, lv)
                    _check_planes(got, planes, f"{w}x{h} q{q} type {ftype}")
        finally:
            s.close()
    finally:
        ctx.set_tuning(**before)


def test_reconstruct_b(ctx, oracle):
    frames = synth.sequence_c1(47, 3, 160, 96, local_range=12)
    fwd, _, bwd = frames
    # A picture midway between its references: bidirectional MBs win.
    rng = np.random.default_rng(3)
    cur = tuple(((f.astype(np.int32) + b + 1) >> 1).astype(np.int16) +
                rng.integers(-1, 2, f.shape).astype(np.int16) for f, b in zip(fwd, bwd))
    h, w = cur[0].shape
    s = ctx.open_stream(w, h, 4)
    for k, f in enumerate((cur, fwd, bwd)):
        s.upload(k, *f)
    s.analyze_b(11, 0, 1, 2, target_ids=[0, 1])
    rec, co, lv = s.results()
    assert (rec["prediction_mode"] == 2).any(), "no bidirectional MB in the sample"
    lv2 = _perturb(lv, 5)
    s.reconstruct(3, lv2)
    want = oracle.reconstruct_frame(2, 11, Picture(*cur), [Picture(*fwd), Picture(*bwd)], rec, lv2)
    _check_planes(s.download(3), want, "B")
    s.close()


def test_reconstruct_wrapping_samples(ctx, oracle):
    """Large residual levels and pictures near the int16 limits: the int16
    stores of the inverse transform wrap exactly as the reference's."""
    rng = np.random.default_rng(77)
    frames = synth.sequence_c1(49, 2, 128, 64, local_range=10)
    big = [tuple(np.clip(p.astype(np.int32) * 120 - 15000, -32768, 32767).astype(np.int16)
                 for p in f) for f in frames]
    h, w = big[0][0].shape
    s = ctx.open_stream(w, h, 3)
    for k, f in enumerate(big):
        s.upload(k, *f)
    s.analyze_p(6, 1, [0])
    rec, co, lv = s.results()
    levels = rng.integers(-4000, 4000, lv.shape).astype(np.int16)
    s.reconstruct(2, levels)
    want = oracle.reconstruct_frame(1, 6, Picture(*big[1]), [Picture(*big[0])], rec, levels)
    _check_planes(s.download(2), want, "wide samples")
    s.close()


def test_closed_loop_reference(ctx, oracle):
    """P frame 1 is reconstructed into slot 3 and serves as frame 2's
    reference, on the device and in the oracle alike."""
    frames = synth.sequence_c1(53, 3, 128, 96, local_range=10)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 4)
    for k, f in enumerate(frames):
        s.upload(k, *f)
    pics = [Picture(*f) for f in frames]
    s.analyze_p(9, 1, [0])
    rec1, _, lv1 = s.results()
    s.reconstruct(3)
    recon = oracle.reconstruct_frame(1, 9, pics[1], [pics[0]], rec1, lv1)
    _check_planes(s.download(3), recon, "frame 1")
    s.analyze_p(9, 2, [3], temporal_mvp=False)
    rec2, co2, lv2 = s.results()
    orec, oco, olv = oracle.analyze_frame(1, 9, pics[2], [Picture(*recon)])
    for f in [n for n in evxa.RECORD_DTYPE.names if not n.startswith("reserved")]:
        assert np.array_equal(rec2[f], orec[f]), f
    assert np.array_equal(co2, oco) and np.array_equal(lv2, olv)
    s.close()


def test_reconstruct_refusals(ctx):
    frames = synth.sequence_c1(59, 2, 64, 48, local_range=6)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 3)
    for k, f in enumerate(frames):
        s.upload(k, *f)
    with pytest.raises(evxa.EvxaError):
        s.reconstruct(2)                       # nothing analysed yet
    s.analyze_p(8, 1, [0])
    with pytest.raises(evxa.EvxaError):
        s.reconstruct(0)                       # the frame's own reference
    with pytest.raises(evxa.EvxaError):
        s.reconstruct(5)                       # no such slot
    s.reconstruct(2)
    s.close()
