"""Frame roles of the hierarchical-B workload (host logic, CPU only)."""
import pytest

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


# ---- the reference encoder's own GOP (evx3enc.cpp:104-239, 279-415) ----

def test_dpb_slot_helpers_match_reference_cases():
    """The exact cases of the reference's DPB slot tests
    (tests/test_codec_roundtrip.cpp:13-31)."""
    assert [gop.assign_ip_dpb_slot(k) for k in range(4)] == [0, 1, 0, 1]
    assert gop.assign_bref_dpb_slot() == 2
    assert gop.assign_b_scratch_slot() == 3
    assert [gop.assign_b_scratch_slot_for_count(k) for k in (1, 2, 3)] == [2, 2, 3]
    # ip_index - 1 at ip_index 0 is uint32 0xFFFFFFFF there: slot 1.
    assert gop.assign_ip_dpb_slot(-1) == 1


def test_internal_quality_mapping():
    # evx3enc.cpp:64-82: identity to 16, then q + (q-16)^2/10, input clipped to 1..31.
    assert [gop.internal_quality(q) for q in range(1, 17)] == list(range(1, 17))
    assert [gop.internal_quality(q) for q in (17, 19, 20, 26, 31)] == [17, 19, 21, 36, 53]
    assert gop.internal_quality(0) == 1 and gop.internal_quality(99) == 53


def _roles(rs):
    return [(r.index, r.kind, r.refs, r.quality, r.slot, r.ref_slots, r.is_reference)
            for r in rs]


def test_reference_schedule_full_gops():
    rs = gop.reference_schedule(9, 8)
    assert _roles(rs) == [
        (0, "I", [], 8, 0, [], True),
        (4, "P", [0], 8, 1, [0], True),
        (2, "B", [0, 4], 9, 2, [0, 1], True),      # reference B: b_count/2 = 1 of 3
        (1, "B", [0, 2], 10, 3, [0, 2], False),
        (3, "B", [2, 4], 10, 3, [2, 1], False),
        (8, "P", [4], 8, 0, [1], True),
        (6, "B", [4, 8], 9, 2, [1, 0], True),
        (5, "B", [4, 6], 10, 3, [1, 2], False),
        (7, "B", [6, 8], 10, 3, [2, 0], False)]
    # P anchors take the previous non-B frame's table (the I frame's, then the P's).
    assert [r.temporal_from for r in rs if r.kind == "P"] == [0, 4]


@pytest.mark.parametrize("n,tail", [
    # flush with one pending frame: a lone P (evx3enc.cpp:389-405)
    (6, [(5, "P", [4], 12, 0, [1], True)]),
    # one B pending: it is the reference B at quality + 1 (ref_b_idx = 1 / 2 = 0)
    (7, [(6, "P", [4], 12, 0, [1], True), (5, "B", [4, 6], 13, 2, [1, 0], True)]),
    # two Bs pending: the later one (ref_b_idx = 2 / 2 = 1) is the reference B
    (8, [(7, "P", [4], 12, 0, [1], True), (6, "B", [4, 7], 13, 2, [1, 0], True),
         (5, "B", [4, 6], 14, 3, [1, 2], False)]),
])
def test_reference_schedule_flush(n, tail):
    rs = _roles(gop.reference_schedule(n, 12))
    assert rs[:5] == _roles(gop.reference_schedule(5, 12))
    assert rs[5:] == tail


@pytest.mark.parametrize("b_count", [0, 1, 2])
def test_reference_schedule_non_hierarchical(b_count):
    """Fewer than 3 Bs configured: every B sits between its two anchors at
    quality + 2 in the scratch slot (evx3enc.cpp:356-376)."""
    rs = gop.reference_schedule(10, 8, b_count=b_count)
    anchors = [r.index for r in rs if r.kind != "B"]
    assert anchors == list(range(0, 10, b_count + 1)) + ([9] if 9 % (b_count + 1) else [])
    for r in rs:
        if r.kind == "B":
            lo = max(a for a in anchors if a < r.index)
            hi = min(a for a in anchors if a > r.index)
            assert r.refs == [lo, hi] and r.quality == 10 and r.slot == 2 and not r.is_reference


@pytest.mark.parametrize("quality,bq", [(29, (30, 31)), (30, (31, 31)), (40, (31, 31))])
def test_reference_schedule_quality_caps(quality, bq):
    rs = gop.reference_schedule(5, quality)
    assert [r.quality for r in rs if r.kind != "B"] == [quality, quality]
    assert (rs[2].quality, rs[3].quality) == bq


def test_reference_schedule_periodic_intra():
    """Every intra_rate-th display frame flushes the pending GOP and restarts
    the anchors at slot 0 (evx3enc.cpp:164-176)."""
    rs = gop.reference_schedule(14, 8, intra_rate=6)
    assert [(r.index, r.kind) for r in rs] == [
        (0, "I"), (4, "P"), (2, "B"), (1, "B"), (3, "B"), (5, "P"),
        (6, "I"), (10, "P"), (8, "B"), (7, "B"), (9, "B"), (11, "P"), (12, "I"), (13, "P")]
    assert [r.slot for r in rs if r.kind != "B"] == [0, 1, 0, 0, 1, 0, 0, 1]


@pytest.mark.parametrize("n,b_count,rate", [(1, 3, 3600), (33, 3, 3600), (40, 3, 7), (23, 2, 3600),
                                            (17, 1, 5), (12, 0, 3600)])
def test_reference_schedule_dpb_holds_every_reference(n, b_count, rate):
    """Decode the schedule into a 4-slot picture buffer: every frame's
    reference slots hold exactly the frames it names, each display frame is
    coded once, and no reference was coded after its user."""
    rs = gop.reference_schedule(n, 8, b_count=b_count, intra_rate=rate)
    assert sorted(r.index for r in rs) == list(range(n))
    dpb = {}
    for r in rs:
        assert [dpb.get(s) for s in r.ref_slots] == r.refs, r
        assert (r.kind == "P") == (len(r.refs) == 1) and (r.kind == "B") == (len(r.refs) == 2)
        dpb[r.slot] = r.index


def test_gop8_at_gop4_equals_reference_on_full_gops():
    """gop8_schedule with gop=4 gives the reference's roles on whole GOPs."""
    for n in (5, 9, 13):
        a = [(r.index, r.kind, r.refs, r.quality) for r in gop.gop8_schedule(n, 12, gop=4)]
        b = [(r.index, r.kind, r.refs, r.quality) for r in gop.reference_schedule(n, 12)]
        assert a == b, n
