"""The oracle against the golden vectors the reference itself wrote
(tests/golden/make_golden.py over oracle/_ref).  CPU only."""
import os

import numpy as np
import pytest

from oracle.oracle import REF_FIELDS, Picture

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def test_lambda_table(oracle):
    g = load("lambda.npz")
    got = np.array([oracle.lam(int(q)) for q in g["quality"]], np.float32).view(np.uint32)
    assert np.array_equal(got, g["bits"])


@pytest.mark.parametrize("tag", ["u8", "wide"])
def test_interp(oracle, tag):
    g = load("interp.npz")
    P = Picture(g[f"{tag}_y"], g[f"{tag}_u"], g[f"{tag}_v"])
    for p, want in zip(g[f"{tag}_params"], g[f"{tag}_out"]):
        assert np.array_equal(oracle.interp(P, *[int(v) for v in p]), want), p


def test_metrics(oracle):
    g = load("metrics.npz")
    for a, b, m in zip(g["a"], g["b"], g["sad_mad_satd"]):
        assert oracle.metrics(a, b) == tuple(int(v) for v in m)
    for x, want in zip(g["had_in"], g["had_out"]):
        assert np.array_equal(oracle.hadamard(x.reshape(4, 4)), want)


def test_dct_quant(oracle):
    g = load("dctquant.npz")
    for k in range(len(g["src"])):
        c = oracle.sub_transform(g["src"][k], g["pred"][k])
        assert np.array_equal(c, g["coeffs"][k]), k
        assert oracle.variance2(c) == g["variance2"][k], k
        for q, want in zip(g["qualities"], g["qp"][k]):
            assert oracle.block_qp(int(q), c) == want, (k, q)
        for q, want in zip(g["qps"], g["levels"][k]):
            assert np.array_equal(oracle.quantize(int(q), c), want), (k, q)


def test_mv_predictor(oracle):
    g = load("mvpred.npz")
    wb = int(g["wb"])
    for t in range(len(g["tables"])):
        for use_prev in (0, 1):
            for idx in range(g["tables"].shape[1]):
                got = oracle.mv_predictor(g["tables"][t], wb, idx,
                                          g["prevs"][t] if use_prev else None)
                assert got == tuple(int(v) for v in g["preds"][t, use_prev, idx]), (t, use_prev, idx)


@pytest.mark.parametrize("tag", ["s64", "s128", "s128c"])
def test_inter_search(oracle, tag):
    g = load("search.npz")
    R = Picture(g[f"{tag}_ref_y"], g[f"{tag}_ref_u"], g[f"{tag}_ref_v"])
    S = Picture(g[f"{tag}_cur_y"], g[f"{tag}_cur_u"], g[f"{tag}_cur_v"])
    copies = 0
    for (x, y, q, pmx, pmy), want in zip(g[f"{tag}_params"], g[f"{tag}_records"]):
        r = oracle.inter_search(S.block(int(x), int(y)), R, int(x), int(y), int(q), int(pmx), int(pmy))
        got = (r.block_type, r.motion_x, r.motion_y, r.sp_enabled, r.sp_amount, r.sp_index, r.sad)
        exp = tuple(int(want[f]) for f in ("block_type", "motion_x", "motion_y", "sp_pred",
                                            "sp_amount", "sp_index", "sad"))
        assert got == exp, (x, y, q, pmx, pmy)
        copies += bool(r.block_type & 4)
    if tag == "s128c":
        assert copies > 0


@pytest.mark.parametrize("tag", ["tex", "tile"])
def test_intra_search(oracle, tag):
    """calculate_intra_prediction against intra.npz (the reference build's
    records for every MB at qualities 0, 8, 13 and 40)."""
    g = load("intra.npz")
    S = Picture(g[f"{tag}_src_y"], g[f"{tag}_src_u"], g[f"{tag}_src_v"])
    P = Picture(g[f"{tag}_pic_y"], g[f"{tag}_pic_u"], g[f"{tag}_pic_v"])
    kinds = set()
    for (x, y, q), want in zip(g[f"{tag}_params"], g[f"{tag}_records"]):
        r = oracle.intra_search(S.block(int(x), int(y)), P, int(x), int(y), int(q))
        got = (r.block_type, r.motion_x, r.motion_y, r.sp_enabled, r.sp_amount, r.sp_index, r.sad)
        exp = tuple(int(want[f]) for f in ("block_type", "motion_x", "motion_y", "sp_pred",
                                            "sp_amount", "sp_index", "sad"))
        assert got == exp, (tag, x, y, q)
        kinds.add(r.block_type)
    if tag == "tile":
        assert {1, 3, 7} <= kinds


def test_bidir(oracle):
    g = load("bidir.npz")
    F = Picture(g["fwd_y"], g["fwd_u"], g["fwd_v"])
    B = Picture(g["bwd_y"], g["bwd_u"], g["bwd_v"])
    S = Picture(g["cur_y"], g["cur_u"], g["cur_v"])
    for (x, y, q, pmx, pmy), want, avg in zip(g["params"], g["records"], g["averages"]):
        x, y, q, pmx, pmy = (int(v) for v in (x, y, q, pmx, pmy))
        src = S.block(x, y)
        f = oracle.inter_search(src, F, x, y, q, pmx, pmy)
        b = oracle.inter_search(src, B, x, y, q, pmx, pmy)
        sad, a = oracle.bidir(src, oracle.predict(F, x, y, f), oracle.predict(B, x, y, b))
        assert sad == want["sad"]
        assert np.array_equal(a, avg)
        assert (f.motion_x, f.motion_y, b.motion_x, b.motion_y) == (
            want["motion_x"], want["motion_y"], want["motion_x_b"], want["motion_y_b"])
        assert (f.sp_enabled, f.sp_index, f.sp_amount, b.sp_enabled, b.sp_index, b.sp_amount) == (
            want["sp_pred"], want["sp_index"], want["sp_amount"], want["sp_pred_b"],
            want["sp_index_b"], want["sp_amount_b"])


def _frame_fixture():
    g = load("frames.npz")
    pics = [Picture(g[f"f{k}_y"], g[f"f{k}_u"], g[f"f{k}_v"]) for k in range(6)]
    return g, pics


def _check(got, g, tag):
    rec, co, lv = got
    for f in REF_FIELDS:
        assert np.array_equal(rec[f], g[f"{tag}_records"][f]), (tag, f)
    assert np.array_equal(co, g[f"{tag}_coeffs"]), tag
    assert np.array_equal(lv, g[f"{tag}_levels"]), tag


def test_frames(oracle):
    g, pics = _frame_fixture()
    r1 = oracle.analyze_frame(1, 8, pics[1], [pics[0]])
    _check(r1, g, "p1")
    r2 = oracle.analyze_frame(1, 8, pics[2], [pics[1], pics[0]], prev=r1[0])
    _check(r2, g, "p2")
    _check(oracle.analyze_frame(1, 12, pics[5], [pics[4], pics[3], pics[2], pics[1]], prev=r2[0]),
           g, "p4")
    _check(oracle.analyze_frame(2, 9, pics[3], [pics[2], pics[4]]), g, "b")


def test_dequantize_and_inverse_transform():
    """quantize.cpp:322-332, 427-449 and transform.cpp:330-433 against the
    reference build's outputs (recon.npz)."""
    from oracle.oracle import Oracle
    o = Oracle()
    g = load("recon.npz")
    for qp, lv, want in zip(g["dq_qp"], g["dq_levels"], g["dq_coeffs"]):
        assert np.array_equal(o.dequantize(int(qp), lv), want)
    for co, pred, want in zip(g["it_coeffs"], g["it_pred"], g["it_out"]):
        assert np.array_equal(o.idct_add(co, pred), want)


@pytest.mark.parametrize("tag,ftype,q,cur,refs", [("p1", 1, 8, 1, [0]), ("p2", 1, 8, 2, [1, 0]),
                                                  ("p4", 1, 12, 5, [4, 3, 2, 1]),
                                                  ("b", 2, 9, 3, [2, 4])])
def test_reconstruct_frames(oracle, tag, ftype, q, cur, refs):
    """decode.cpp:123-433 composed from the reference's functions, on the
    analysed frames of frames.npz with their own and perturbed levels."""
    g, pics = _frame_fixture()
    r = load("recon.npz")
    rec = g[f"{tag}_records"]
    for name, levels in (("own", g[f"{tag}_levels"]), ("perturbed", r[f"{tag}_levels_perturbed"])):
        y, u, v = oracle.reconstruct_frame(ftype, q, pics[cur], [pics[k] for k in refs], rec, levels)
        assert np.array_equal(y, r[f"{tag}_{name}_y"]), (tag, name)
        assert np.array_equal(u, r[f"{tag}_{name}_u"]), (tag, name)
        assert np.array_equal(v, r[f"{tag}_{name}_v"]), (tag, name)


def test_deblock_lines(oracle):
    """deblock.cpp:91-165 on random lines against the reference build
    (deblock.npz), extreme samples included."""
    g = load("deblock.npz")
    for v, qp, s, lu, want in zip(g["line_in"], g["line_qp"], g["line_strength"], g["line_luma"],
                                  g["line_out"]):
        assert np.array_equal(oracle.deblock_line(v, int(qp), int(s), int(lu)), want), (v, qp, s, lu)


@pytest.mark.parametrize("tag", ["p1", "p4", "b"])
def test_deblock_reconstructed_frames(oracle, tag):
    """deblock.cpp:222-306 on the reconstructions of recon.npz with the
    records of their analysis."""
    g, _ = _frame_fixture()
    r, d = load("recon.npz"), load("deblock.npz")
    out = oracle.deblock_frame(g[f"{tag}_records"], *(r[f"{tag}_perturbed_{c}"] for c in "yuv"))
    for c, plane in zip("yuv", out):
        assert np.array_equal(plane, d[f"{tag}_{c}"]), (tag, c)


@pytest.mark.parametrize("tag", ["blocky", "extreme"])
def test_deblock_synthetic_frames(oracle, tag):
    """Random block types and q_index over a blocky picture, and over
    samples near the int16 limits whose edge differences wrap."""
    d = load("deblock.npz")
    out = oracle.deblock_frame(d[f"{tag}_records"], *(d[f"{tag}_in_{c}"] for c in "yuv"))
    for c, plane in zip("yuv", out):
        assert np.array_equal(plane, d[f"{tag}_{c}"]), (tag, c)


def _sao_same(a, b):
    from oracle.oracle import SAO_DTYPE
    return all(np.array_equal(a[f], b[f]) for f in SAO_DTYPE.names if "reserved" not in f)


def test_sao_filter_table(oracle):
    """sao.cpp:98-148 under a random table against the reference build
    (sao.npz)."""
    g = load("sao.npz")
    out = oracle.sao_filter_frame(g["filter_table"], *(g[f"filter_in_{c}"] for c in "yuv"))
    for c, plane in zip("yuv", out):
        assert np.array_equal(plane, g[f"filter_{c}"]), c


@pytest.mark.parametrize("tag,cur,q", [("p1", 1, 8), ("p4", 5, 12), ("b", 3, 9)])
def test_sao_encode_deblocked_frames(oracle, tag, cur, q):
    """sao.cpp:378-590 on the deblocked reconstructions of deblock.npz
    against their source frames."""
    f, d, g = load("frames.npz"), load("deblock.npz"), load("sao.npz")
    table, *planes = oracle.sao_encode_frame([f[f"f{cur}_{c}"] for c in "yuv"],
                                             [d[f"{tag}_{c}"] for c in "yuv"], oracle.lam(q))
    assert _sao_same(table, g[f"{tag}_table"])
    for c, plane in zip("yuv", planes):
        assert np.array_equal(plane, g[f"{tag}_{c}"]), c


@pytest.mark.parametrize("q", [6, 20])
def test_sao_encode_mixed_content(oracle, q):
    """Edge, band and no offsets, and merges left and up."""
    g = load("sao.npz")
    table, *planes = oracle.sao_encode_frame([g[f"mix_src_{c}"] for c in "yuv"],
                                             [g[f"mix_rec_{c}"] for c in "yuv"], oracle.lam(q))
    assert _sao_same(table, g[f"mix{q}_table"])
    for c, plane in zip("yuv", planes):
        assert np.array_equal(plane, g[f"mix{q}_{c}"]), c
