"""The in-loop encode on the device (evxa_encode_frame: encode_slice with the
intra-frame block search, the inter search, directional intra modes, RDO,
trellis and the reconstruction into the picture buffer) against the oracle
and against the reference's own encode_slice vectors (encode.npz).

Every case is a chain I -> P -> B whose reconstructions serve as the next
frames' references, each frame reconstructed over the stale content of its
picture-buffer slot.  Compared bit for bit: every descriptor field, the
coefficients, the post-trellis levels and the reconstructed picture."""
import os

import numpy as np
import pytest

from chicago_amd import evxa, synth
from oracle.oracle import Picture

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = [n for n in evxa.RECORD_DTYPE.names if not n.startswith("reserved")]
CHAINS = ("nat8", "nat20", "tile1", "tile8", "tile20", "flat8")


def gpu_chain(ctx, frames, stale, quality, ids=(5, 6), slots=None):
    """I, P (temporal table = the I records) and B of `frames` on the
    device: (records, coefficients, levels, planes) per frame type."""
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 6)
    try:
        for k in range(3):
            s.upload(k, *frames[k])
        for k in (3, 4, 5):
            s.upload(k, *stale)
        out = {}
        s.encode_frame(evxa.FRAME_I, quality, 0, 3)
        out["I"] = (*s.results(), s.download(3))
        s.encode_frame(evxa.FRAME_P, quality, 1, 4, [3], target_ids=[ids[0]])
        out["P"] = (*s.results(), s.download(4))
        s.encode_frame(evxa.FRAME_B, quality, 2, 5, [3, 4], target_ids=list(ids))
        out["B"] = (*s.results(), s.download(5))
        return out
    finally:
        s.close()


def oracle_chain(oracle, frames, stale, quality, ids=(5, 6)):
    src = [Picture(*f) for f in frames]
    ri, li, pi, ci = oracle.encode_frame_coeffs(0, quality, src[0], [], stale)
    rp, lp, pp, cp = oracle.encode_frame_coeffs(1, quality, src[1], [Picture(*pi)], stale,
                                                prev=ri, target_ids=[ids[0]])
    rb, lb, pb, cb = oracle.encode_frame_coeffs(2, quality, src[2],
                                                [Picture(*pi), Picture(*pp)], stale,
                                                target_ids=list(ids))
    return {"I": (ri, ci, li, pi), "P": (rp, cp, lp, pp), "B": (rb, cb, lb, pb)}


def assert_chain_equal(got, want, tag):
    for t in "IPB":
        rec, co, lv, pic = got[t]
        wrec, wco, wlv, wpic = want[t]
        for f in FIELDS:
            bad = np.nonzero(rec[f] != wrec[f])[0]
            assert bad.size == 0, (tag, t, f, bad[:6], rec[f][bad[:6]], wrec[f][bad[:6]])
        assert np.array_equal(lv, wlv), (tag, t, "levels",
                                         np.nonzero((lv != wlv).any(axis=1))[0][:6])
        assert np.array_equal(co, wco), (tag, t, "coefficients")
        for c, a, b in zip("yuv", pic, wpic):
            bad = np.argwhere(a != b)
            assert bad.size == 0, (tag, t, c, bad[:4])


def planes(g, key):
    return tuple(g[f"{key}_{c}"] for c in "yuv")


@pytest.mark.parametrize("key", CHAINS)
def test_encode_golden(ctx, oracle, key):
    """The reference's own encode_slice vectors: descriptors, levels and
    pictures of I, P and B; every field and coefficient also vs the oracle."""
    print(f"encode golden {key}", flush=True)
    g = np.load(os.path.join(GOLD, "encode.npz"))
    stale = planes(g, f"{key}_stale")
    frames = [planes(g, f"{key}_src{i}") for i in range(3)]
    q = int(key.lstrip("abcdefghijklmnopqrstuvwxyz"))
    got = gpu_chain(ctx, frames, stale, q)
    fields = tuple(g["fields"])
    for t in "IPB":
        rec, _, lv, pic = got[t]
        have = np.stack([rec[f].astype(np.int64) for f in fields], axis=-1)
        want = g[f"{key}_{t}_fields"]
        bad = np.nonzero((have != want).any(axis=1))[0]
        assert bad.size == 0, (key, t, bad[:5], have[bad[:3]], want[bad[:3]])
        assert np.array_equal(lv, g[f"{key}_{t}_levels"]), (key, t)
        for c, p in zip("yuv", pic):
            assert np.array_equal(p, g[f"{key}_{t}_{c}"]), (key, t, c)
    assert_chain_equal(got, oracle_chain(oracle, frames, stale, q), key)


def random_pictures(rng, w, h, kind):
    if kind == "u8":
        return tuple(rng.integers(0, 256, s).astype(np.int16)
                     for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2)))
    if kind == "wide":
        return tuple(rng.integers(-3000, 3000, s).astype(np.int16)
                     for s in ((h, w), (h // 2, w // 2), (h // 2, w // 2)))
    return synth.texture_frame(rng, w, h)


@pytest.mark.parametrize("kind,quality,w,h", [
    ("tex", 4, 128, 96), ("tex", 12, 128, 96), ("tex", 24, 128, 96),
    ("u8", 8, 128, 96), ("wide", 12, 128, 96),
    ("tex", 0, 128, 96), ("wide", 40, 128, 96), ("u8", 63, 128, 96),
    ("tex", 8, 48, 48), ("tex", 8, 320, 176), ("tex", 20, 256, 64)])
def test_encode_vs_oracle(ctx, oracle, kind, quality, w, h):
    """Fresh chains (natural motion, noise, 8-bit noise, values outside 8
    bits; the smallest frame and frames with many MBs in flight)."""
    print(f"encode {kind} q{quality} {w}x{h}", flush=True)
    rng = np.random.default_rng(quality * 31 + w + h + len(kind))
    if kind == "tex":
        frames = synth.sequence_c1(int(rng.integers(1 << 30)), 3, w, h, local_range=10)
    else:
        base = random_pictures(rng, w, h, kind)
        frames = [base] + [tuple(np.roll(p, (k, 2 * k), axis=(0, 1)) for p in base)
                           for k in (1, 2)]
    stale = random_pictures(rng, w, h, "u8")
    got = gpu_chain(ctx, frames, stale, quality, ids=(3, 1))
    assert_chain_equal(got, oracle_chain(oracle, frames, stale, quality, ids=(3, 1)),
                       (kind, quality, w, h))


def test_encode_b_one_picture_both_ways(ctx, oracle):
    """A B frame whose forward and backward references are one slot under one
    target id, against the oracle (which matches the reference build there)."""
    print("encode B, one reference picture both ways", flush=True)
    frames = synth.sequence_c1(77, 3, 96, 64, local_range=10)
    rng = np.random.default_rng(78)
    stale = random_pictures(rng, 96, 64, "u8")
    s = ctx.open_stream(96, 64, 4)
    try:
        s.upload(0, *frames[0])
        s.upload(1, *frames[2])
        for k in (2, 3):
            s.upload(k, *stale)
        s.encode_frame(evxa.FRAME_I, 12, 0, 2)
        gi = (*s.results(), s.download(2))
        s.encode_frame(evxa.FRAME_B, 12, 1, 3, [2, 2], target_ids=[2, 2])
        gb = (*s.results(), s.download(3))
    finally:
        s.close()
    ri, li, pi, ci = oracle.encode_frame_coeffs(0, 12, Picture(*frames[0]), [], stale)
    rb, lb, pb, cb = oracle.encode_frame_coeffs(2, 12, Picture(*frames[2]),
                                                [Picture(*pi), Picture(*pi)], stale,
                                                target_ids=[2, 2])
    for tag, got, want in (("I", gi, (ri, ci, li, pi)), ("B", gb, (rb, cb, lb, pb))):
        rec, co, lv, pic = got
        for f in FIELDS:
            assert np.array_equal(rec[f], want[0][f]), (tag, f)
        assert np.array_equal(co, want[1]) and np.array_equal(lv, want[2]), tag
        for a, b in zip(pic, want[3]):
            assert np.array_equal(a, b), tag


def test_encode_720p_vs_oracle(ctx, oracle):
    """1280x720 (the CPU-runnable pair's size): I then P frames, every MB."""
    print("encode 1280x720", flush=True)
    frames = synth.sequence_c1(71, 3, 1280, 720)
    stale = tuple(np.zeros_like(p) for p in frames[0])
    got = gpu_chain(ctx, frames, stale, 8)
    assert_chain_equal(got, oracle_chain(oracle, frames, stale, 8), "720p")


def test_encode_1080p_vs_oracle(ctx, oracle):
    """1920x1088 (the 1080p frames of the bench), I, P and B: every field,
    coefficient, level and sample of all 8160 MBs."""
    print("encode 1920x1088", flush=True)
    frames = synth.sequence_c1(72, 3, 1920, 1088)
    stale = tuple(np.zeros_like(p) for p in frames[0])
    got = gpu_chain(ctx, frames, stale, 8)
    assert_chain_equal(got, oracle_chain(oracle, frames, stale, 8), "1080p")


def test_encode_refusals(ctx):
    """Bad jobs are refused before anything runs: the destination is the
    source or a reference, a frame type, a quality, a frame below 48x48."""
    s = ctx.open_stream(64, 48, 3)
    try:
        with pytest.raises(evxa.EvxaError):
            s.encode_frame(evxa.FRAME_I, 8, 0, 0)
        with pytest.raises(evxa.EvxaError):
            s.encode_frame(evxa.FRAME_P, 8, 0, 1, [1])
        with pytest.raises(evxa.EvxaError):
            s.encode(evxa.encode_desc(3, 8, 0, 1))
        with pytest.raises(evxa.EvxaError):
            s.encode_frame(evxa.FRAME_I, 64, 0, 1)
        # Two reference pictures under one target id (the decoder's slot).
        with pytest.raises(evxa.EvxaError):
            s.encode_frame(evxa.FRAME_B, 8, 0, 1, [2, 0], target_ids=[4, 4])
    finally:
        s.close()
    # evxa_reconstruct after an encode (which reconstructs its frame itself)
    s = ctx.open_stream(64, 48, 3)
    try:
        s.encode_frame(evxa.FRAME_I, 8, 0, 1)
        with pytest.raises(evxa.EvxaError):
            s.reconstruct(2)
        s.deblock(1)                     # the encode's own descriptors
        s.wait()
    finally:
        s.close()
    small = ctx.open_stream(32, 32, 2)
    try:
        with pytest.raises(evxa.EvxaError):
            small.encode_frame(evxa.FRAME_I, 8, 0, 1)
    finally:
        small.close()
    # evxa_encode_batch: one stream twice in a batch, a stream of another
    # context.
    s = ctx.open_stream(64, 48, 3)
    other = evxa.Context(0)
    t = other.open_stream(64, 48, 3)
    try:
        d = evxa.encode_desc(evxa.FRAME_I, 8, 0, 1)
        with pytest.raises(evxa.EvxaError):
            ctx.encode_batch([s, s], [d, d])
        with pytest.raises(evxa.EvxaError):
            ctx.encode_batch([s, t], [d, d])
        ctx.encode_batch([s], [d])
        s.wait()
    finally:
        t.close()
        other.close()
        s.close()


def test_encode_batch_matches_alone(ctx, oracle):
    """evxa_encode_batch: frames of streams of uneven sizes (and types) coded
    in one launch give each stream's outputs of coding it alone."""
    print("encode batch", flush=True)
    sizes = [(128, 96), (64, 48), (320, 176), (96, 64)]
    seqs = [synth.sequence_c1(100 + k, 3, w, h, local_range=10) for k, (w, h) in enumerate(sizes)]
    stale = [tuple(np.full_like(p, 17 * k) for p in seqs[k][0]) for k in range(len(sizes))]
    streams = []
    try:
        for k, (w, h) in enumerate(sizes):
            s = ctx.open_stream(w, h, 6)
            for f in range(3):
                s.upload(f, *seqs[k][f])
            for f in (3, 4, 5):
                s.upload(f, *stale[k])
            streams.append(s)
        steps = [
            [evxa.encode_desc(evxa.FRAME_I, 8, 0, 3) for _ in sizes],
            [evxa.encode_desc(evxa.FRAME_P, 8, 1, 4, [3], target_ids=[5]) for _ in sizes],
            [evxa.encode_desc(evxa.FRAME_B, 8, 2, 5, [3, 4], target_ids=[5, 6]) for _ in sizes],
        ]
        got = [{} for _ in sizes]
        for t, descs in zip("IPB", steps):
            ctx.encode_batch(streams, descs)
            for k, s in enumerate(streams):
                got[k][t] = (*s.results(), s.download(3 + "IPB".index(t)))
        for k in range(len(sizes)):
            alone = gpu_chain(ctx, seqs[k], stale[k], 8)
            assert_chain_equal(got[k], alone, ("batch vs alone", k))
        assert_chain_equal(got[2], oracle_chain(oracle, seqs[2], stale[2], 8), "batch vs oracle")
    finally:
        for s in streams:
            s.close()
