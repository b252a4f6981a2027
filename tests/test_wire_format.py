"""Records and SAO tables against the reference's wire format: the unified
block table (serialize.cpp:184-314) and the SAO table (serialize.cpp:564-622)
arithmetic-coded by the reference's own serializer, from the oracle's records
converted to evx_block_desc, byte for byte equal to the serialization of the
reference analysis's own descriptors (never converted).  This pins that a
record carries every field the bitstream reads, including what the MV
differences need (compute_mv_predictor, serialize.cpp:244).  CPU only;
skipped where the reference build is absent."""
import numpy as np
import pytest

from chicago_amd import synth
from oracle.oracle import Picture, reference_available

pytestmark = pytest.mark.skipif(not reference_available(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    from oracle.oracle import Reference
    return Reference()


@pytest.fixture(scope="module")
def pics():
    return [Picture(*f) for f in synth.sequence_c1(5200, 6, 176, 112, local_range=16)]


@pytest.mark.parametrize("ftype,nrefs,quality", [(1, 1, 6), (1, 2, 10), (1, 4, 14), (2, 2, 9)])
def test_block_table_bytes(oracle, ref, pics, ftype, nrefs, quality):
    if ftype == 1:
        cur, refs = pics[5], [pics[4 - k] for k in range(nrefs)]
    else:
        cur, refs = pics[3], [pics[2], pics[5]]
    rec, _, _ = oracle.analyze_frame(ftype, quality, cur, refs)
    assert len(np.unique(rec["block_type"])) >= 2
    want = ref.analyze_frame_serialized(ftype, quality, cur, refs)
    got = ref.serialize_records(rec, cur.width, cur.height, ftype)
    assert got == want, (len(got), len(want))


def test_block_table_bytes_temporal(oracle, ref, pics):
    """With the previous frame's table as the temporal MV predictor."""
    prev, _, _ = oracle.analyze_frame(1, 10, pics[4], [pics[3], pics[2]])
    rec, _, _ = oracle.analyze_frame(1, 10, pics[5], [pics[4], pics[3]], prev=prev)
    want = ref.analyze_frame_serialized(1, 10, pics[5], [pics[4], pics[3]], prev=prev)
    got = ref.serialize_records(rec, pics[5].width, pics[5].height, 1, prev=prev)
    assert got == want


def test_block_table_bytes_custom_target_ids(oracle, ref, pics):
    """Records carry the caller's target ids; converted back to reference
    indexes they serialize as the reference's own."""
    refs = [pics[4], pics[3], pics[2]]
    ids = [7, 3, 9]
    rec, _, _ = oracle.analyze_frame(1, 12, pics[5], refs, target_ids=ids)
    want = ref.analyze_frame_serialized(1, 12, pics[5], refs, target_ids=ids)
    assert ref.serialize_records(rec, pics[5].width, pics[5].height, 1, target_ids=ids) == want


@pytest.mark.parametrize("q", [6, 20])
def test_sao_table_bytes(oracle, ref, q):
    rng = np.random.default_rng(q)
    w, h = 176, 144
    src = synth.blocky_picture(44 + q, w, h, amp=3)
    rec = tuple((p.astype(np.int32) + rng.integers(-4, 5, p.shape)).astype(np.int16) for p in src)
    mine, *_ = oracle.sao_encode_frame(src, rec, oracle.lam(q))
    theirs, *_ = ref.sao_encode_frame(src, rec, ref.lam(q))
    assert mine["merge_left"].any()
    assert ref.serialize_sao(mine, w, h) == ref.serialize_sao(theirs, w, h)
