"""Frame roles of the hierarchical-B workload (host logic, CPU only)."""
from chicago_amd import gop, synth


def test_gop8_pyramid_roles():
    roles = gop.gop8_schedule(17, 12)
    by_index = {r.index: r for r in roles}
    assert [r.index for r in roles] == [0, 8, 4, 2, 1, 3, 6, 5, 7, 16, 12, 10, 9, 11, 14, 13, 15]
    # B4 <- (0, 8); B2 <- (0, 4); B6 <- (4, 8); B1 <- (0, 2); B3 <- (2, 4); B5 <- (4, 6); B7 <- (6, 8)
    assert {k: by_index[k].refs for k in (4, 2, 6, 1, 3, 5, 7)} == {
        4: [0, 8], 2: [0, 4], 6: [4, 8], 1: [0, 2], 3: [2, 4], 5: [4, 6], 7: [6, 8]}
    assert by_index[8].kind == "P" and by_index[8].refs == [0] and not by_index[8].temporal_mvp
    assert by_index[16].temporal_mvp and by_index[16].refs == [8]
    # Reference Bs (later frames predict from them) at +1, the rest at +2
    # (evx3enc.cpp:318, 337).
    assert {k: by_index[k].quality for k in (4, 2, 6)} == {4: 13, 2: 13, 6: 13}
    assert {k: by_index[k].quality for k in (1, 3, 5, 7)} == {1: 14, 3: 14, 5: 14, 7: 14}
    assert {k: by_index[k].level for k in (4, 2, 1)} == {4: 1, 2: 2, 1: 3}


def test_gop8_decode_order_is_causal():
    for n in (2, 5, 9, 17, 33, 34):
        roles = gop.gop8_schedule(n, 30)
        seen = set()
        for r in roles:
            assert all(x in seen for x in r.refs), (n, r)
            assert r.quality <= 31
            seen.add(r.index)
        assert seen == set(range(n))


def test_sequence_c2_shape_and_range():
    frames = synth.sequence_c2(3, 3, 96, 72)
    assert len(frames) == 3
    for y, u, v in frames:
        assert y.shape == (80, 96) and u.shape == (40, 48) and v.shape == (40, 48)
        assert y.min() >= 0 and y.max() <= 255
        assert not y[72:].any() and not u[36:].any()
