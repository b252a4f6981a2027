"""The device's records and SAO tables through the reference's serializer:
byte for byte the serialization of the reference analysis's own descriptors
and SAO decisions (serialize.cpp:184-314, 564-622).  Skipped where the
reference build is absent."""
import numpy as np
import pytest

from chicago_amd import synth
from oracle.oracle import Picture, reference_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not reference_available(), reason="oracle/_ref not built")]


@pytest.fixture(scope="module")
def ref():
    from oracle.oracle import Reference
    return Reference()


@pytest.mark.parametrize("ftype", [1, 2])
def test_device_records_serialize_as_reference(ctx, ref, ftype):
    frames = synth.sequence_c1(5300 + ftype, 6, 320, 192, local_range=16)
    h, w = frames[0][0].shape
    s = ctx.open_stream(w, h, 6)
    for k, f in enumerate(frames):
        s.upload(k, *f)
    pics = [Picture(*f) for f in frames]
    if ftype == 1:
        s.analyze_p(10, 5, [4, 3, 2, 1])
        args = (1, 10, pics[5], [pics[4], pics[3], pics[2], pics[1]])
    else:
        s.analyze_b(10, 3, 2, 5, target_ids=[0, 1])
        args = (2, 10, pics[3], [pics[2], pics[5]])
    rec, _, _ = s.results()
    assert ref.serialize_records(rec, w, h, ftype) == ref.analyze_frame_serialized(*args)
    s.close()


def test_device_sao_table_serializes_as_reference(ctx, ref):
    rng = np.random.default_rng(8)
    w, h = 320, 192
    src = synth.blocky_picture(61, w, h, amp=3)
    rec = tuple((p.astype(np.int32) + rng.integers(-4, 5, p.shape)).astype(np.int16) for p in src)
    s = ctx.open_stream(w, h, 2)
    s.upload(0, *src)
    s.upload(1, *rec)
    table = s.sao(1, 0, 12)
    theirs, *_ = ref.sao_encode_frame(src, rec, ref.lam(12))
    assert ref.serialize_sao(table, w, h) == ref.serialize_sao(theirs, w, h)
    s.close()
