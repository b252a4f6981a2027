"""The HIP path against the golden vectors the reference's own build wrote
(tests/golden/make_golden.py over oracle/_ref), on the GPU box itself: the
per-block entries (interpolation, DCT + QP + quantisation, searches), whole
analysed P and B frames in both ways of resolving the MV predictor,
reconstruction with the frames' own and perturbed levels, and SAO.  These
pin the device's results to the reference directly, not through the oracle
built on the box.  The intra block search (intra.npz) is pinned the same way."""
import os

import numpy as np
import pytest

from chicago_amd import evxa
from oracle.oracle import REF_FIELDS, SAO_DTYPE

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = [pytest.param(evxa.MODE_WAVEFRONT, id="wavefront"), pytest.param(evxa.MODE_SWEEP, id="sweep")]


def load(name):
    return np.load(os.path.join(GOLD, name))


def _planes(g, tag):
    return tuple(g[f"{tag}_{c}"] for c in "yuv")


def _stream(ctx, pics):
    h, w = pics[0][0].shape
    s = ctx.open_stream(w, h, len(pics) + 1)
    for k, p in enumerate(pics):
        s.upload(k, *p)
    return s


def _same_planes(got, want, what):
    for c, a, b in zip("yuv", got, want):
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{what}: plane {c} differs at {bad[:4].tolist()}"


@pytest.fixture
def set_mode(ctx):
    before = ctx.tuning()[0]
    yield lambda mode: ctx.set_tuning(mode=mode)
    ctx.set_tuning(**before)


@pytest.mark.parametrize("tag", ["u8", "wide"])
def test_interp(ctx, tag):
    g = load("interp.npz")
    s = _stream(ctx, [_planes(g, tag)])
    p = g[f"{tag}_params"]
    got = s.interpolate_subpixel_6tap(0, p[:, 0], p[:, 1], p[:, 2], p[:, 3], p[:, 4])
    bad = np.nonzero((got != g[f"{tag}_out"]).any(1))[0]
    assert bad.size == 0, f"{bad.size} blocks differ, first params {p[bad[:3]].tolist()}"
    s.close()


def test_dct_quant(ctx):
    """Coefficients and variance2 for every pair; QP for every quality; the
    levels wherever that QP is one of the fixture's quantiser steps."""
    g = load("dctquant.npz")
    n = len(g["src"])
    compared = 0
    for j, q in enumerate(g["qualities"]):
        co, lv, qp, var = ctx.dct_quant_blocks(g["src"], g["pred"], np.full(n, q, np.int32))
        assert np.array_equal(co, g["coeffs"]), q
        assert np.array_equal(var, g["variance2"]), q
        assert np.array_equal(qp, g["qp"][:, j]), q
        steps = list(g["qps"])
        for k in range(n):
            if int(qp[k]) in steps:
                assert np.array_equal(lv[k], g["levels"][k, steps.index(int(qp[k]))]), (k, q)
                compared += 1
    assert compared > 0


def _search_fields(r):
    return (int(r["block_type"]), int(r["motion_x"]), int(r["motion_y"]), int(r["sp_enabled"]),
            int(r["sp_amount"]), int(r["sp_index"]), int(r["sad"]))


def _record_fields(r):
    return tuple(int(r[f]) for f in ("block_type", "motion_x", "motion_y", "sp_pred",
                                     "sp_amount", "sp_index", "sad"))


@pytest.mark.parametrize("tag", ["s64", "s128", "s128c"])
def test_search(ctx, tag):
    g = load("search.npz")
    s = _stream(ctx, [_planes(g, f"{tag}_cur"), _planes(g, f"{tag}_ref")])
    params = g[f"{tag}_params"]
    want = g[f"{tag}_records"].view(evxa.RECORD_DTYPE)
    for q in np.unique(params[:, 2]):
        sel = np.nonzero(params[:, 2] == q)[0]
        p = params[sel]
        got = s.perform_inter_search(0, 1, int(q), p[:, 0], p[:, 1], p[:, 3], p[:, 4])
        for k, i in enumerate(sel):
            assert _search_fields(got[k]) == _record_fields(want[i]), (tag, params[i].tolist())
    s.close()


@pytest.mark.parametrize("tag", ["tex", "tile"])
def test_intra_search(ctx, tag):
    """evxa_intra_search_blocks against intra.npz (the reference build's
    calculate_intra_prediction records)."""
    g = load("intra.npz")
    s = _stream(ctx, [_planes(g, f"{tag}_src"), _planes(g, f"{tag}_pic")])
    params = g[f"{tag}_params"]
    want = g[f"{tag}_records"].view(evxa.RECORD_DTYPE)
    for q in np.unique(params[:, 2]):
        sel = np.nonzero(params[:, 2] == q)[0]
        p = params[sel]
        got = s.intra_search(0, 1, int(q), p[:, 0], p[:, 1])
        for k, i in enumerate(sel):
            assert _search_fields(got[k]) == _record_fields(want[i]), (tag, params[i].tolist())
    s.close()


def test_bidir_searches(ctx):
    """The forward and backward searches of bidir.npz's blocks (their
    motion and sub-pel fields; the average is the frame tests' business)."""
    g = load("bidir.npz")
    s = _stream(ctx, [_planes(g, "cur"), _planes(g, "fwd"), _planes(g, "bwd")])
    params = g["params"]
    want = g["records"].view(evxa.RECORD_DTYPE)
    for q in np.unique(params[:, 2]):
        sel = np.nonzero(params[:, 2] == q)[0]
        p = params[sel]
        f = s.perform_inter_search(0, 1, int(q), p[:, 0], p[:, 1], p[:, 3], p[:, 4])
        b = s.perform_inter_search(0, 2, int(q), p[:, 0], p[:, 1], p[:, 3], p[:, 4])
        for k, i in enumerate(sel):
            w = want[i]
            assert (int(f["motion_x"][k]), int(f["motion_y"][k]), int(b["motion_x"][k]),
                    int(b["motion_y"][k])) == (int(w["motion_x"]), int(w["motion_y"]),
                                               int(w["motion_x_b"]), int(w["motion_y_b"])), i
            assert (int(f["sp_enabled"][k]), int(f["sp_index"][k]), int(f["sp_amount"][k]),
                    int(b["sp_enabled"][k]), int(b["sp_index"][k]), int(b["sp_amount"][k])) == (
                int(w["sp_pred"]), int(w["sp_index"]), int(w["sp_amount"]), int(w["sp_pred_b"]),
                int(w["sp_index_b"]), int(w["sp_amount_b"])), i
    s.close()


@pytest.mark.parametrize("mode", MODES)
def test_frames_and_reconstruction(ctx, set_mode, mode):
    """frames.npz in decode order -- P with 1 reference, P with 2 and the
    temporal table, P with 4 and the temporal table, B -- then each frame
    rebuilt with its own and with perturbed levels (recon.npz)."""
    set_mode(mode)
    g, r = load("frames.npz"), load("recon.npz")
    pics = [_planes(g, f"f{k}") for k in range(6)]
    s = _stream(ctx, pics)
    spare = len(pics)
    for tag, ftype, q, cur, refs in (("p1", 1, 8, 1, [0]), ("p2", 1, 8, 2, [1, 0]),
                                     ("p4", 1, 12, 5, [4, 3, 2, 1]), ("b", 2, 9, 3, [2, 4])):
        if ftype == 1:
            s.analyze_p(q, cur, refs)
        else:
            s.analyze_b(q, cur, refs[0], refs[1])
        rec, co, lv = s.results()
        want = g[f"{tag}_records"].view(evxa.RECORD_DTYPE)
        for f in REF_FIELDS:
            assert np.array_equal(rec[f], want[f]), (tag, f)
        assert np.array_equal(co, g[f"{tag}_coeffs"]), tag
        assert np.array_equal(lv, g[f"{tag}_levels"]), tag
        for name, levels in (("own", None), ("perturbed", r[f"{tag}_levels_perturbed"])):
            s.reconstruct(spare, levels)
            _same_planes(s.download(spare), _planes(r, f"{tag}_{name}"), f"{tag} {name}")
    s.close()


def _same_table(got, want, what):
    for f in SAO_DTYPE.names:
        if "reserved" in f:
            continue
        bad = np.argwhere(np.asarray(got[f] != want[f]).reshape(len(got), -1).any(1))
        assert bad.size == 0, f"{what}: field {f} differs at MBs {bad[:6].ravel().tolist()}"


def test_sao_filter_table(ctx):
    g = load("sao.npz")
    s = _stream(ctx, [_planes(g, "filter_in")])
    s.sao_filter(0, g["filter_table"].view(SAO_DTYPE))
    _same_planes(s.download(0), _planes(g, "filter"), "filter")
    s.close()


@pytest.mark.parametrize("tag,cur,q", [("p1", 1, 8), ("p4", 5, 12), ("b", 3, 9)])
def test_sao_decisions_deblocked_frames(ctx, tag, cur, q):
    f, d, g = load("frames.npz"), load("deblock.npz"), load("sao.npz")
    s = _stream(ctx, [_planes(d, tag), _planes(f, f"f{cur}")])
    table = s.sao(0, 1, q)
    _same_table(table, g[f"{tag}_table"].view(SAO_DTYPE), tag)
    _same_planes(s.download(0), _planes(g, tag), tag)
    s.close()


@pytest.mark.parametrize("q", [6, 20])
def test_sao_decisions_mixed_content(ctx, q):
    g = load("sao.npz")
    s = _stream(ctx, [_planes(g, "mix_rec"), _planes(g, "mix_src")])
    table = s.sao(0, 1, q)
    _same_table(table, g[f"mix{q}_table"].view(SAO_DTYPE), f"mix q{q}")
    _same_planes(s.download(0), _planes(g, f"mix{q}"), f"mix q{q}")
    s.close()
