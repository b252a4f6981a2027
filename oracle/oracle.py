"""ctypes bindings for the oracle (liboracle.so) and the reference build
(_ref/libevxref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker of the HIP path.  Planes are numpy
int16 arrays; MB-order blocks are int16[384] (16x16 Y, 8x8 U, 8x8 V).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libevxref.so")

I16P = C.POINTER(C.c_int16)
I32P = C.POINTER(C.c_int32)


class Planes(C.Structure):
    _fields_ = [("y", I16P), ("u", I16P), ("v", I16P), ("width", C.c_int32),
                ("height", C.c_int32), ("stride_y", C.c_int32), ("stride_c", C.c_int32)]


class SearchResult(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("best_x", "best_y", "sad", "mad", "ssd")] + [
        ("cost_bits", C.c_uint32)] + [(n, C.c_int32) for n in (
            "sp_enabled", "sp_amount", "sp_index", "block_type", "motion_x", "motion_y",
            "n_int", "n_axial", "n_diag")]


# Record layout shared with evxa_block_record (64 bytes).
RECORD_DTYPE = np.dtype([
    ("block_type", "u1"), ("prediction_mode", "u1"), ("prediction_target", "u1"),
    ("prediction_target_b", "u1"), ("motion_x", "<i2"), ("motion_y", "<i2"),
    ("motion_x_b", "<i2"), ("motion_y_b", "<i2"), ("sp_pred", "u1"), ("sp_amount", "u1"),
    ("sp_index", "u1"), ("sp_pred_b", "u1"), ("sp_amount_b", "u1"), ("sp_index_b", "u1"),
    ("q_index", "u1"), ("intra_mode", "u1"), ("variance", "<i2"), ("reserved1", "<i2"),
    ("sad", "<i4"), ("mad", "<i4"), ("cost_bits", "<u4"), ("sad_b", "<i4"), ("mad_b", "<i4"),
    ("cost_b_bits", "<u4"), ("n_int", "<u4"), ("n_axial", "<u2"), ("n_diag", "<u2"),
    ("flags", "<u4"), ("reserved2", "<u4")])
assert RECORD_DTYPE.itemsize == 64

# evo_sao_info: one per MB (sao.h:62-74).
SAO_DTYPE = np.dtype([("merge_left", "u1"), ("merge_above", "u1"), ("reserved", "u1", 2),
                      ("luma_type", "u1"), ("luma_eo_class", "u1"), ("luma_band_position", "u1"),
                      ("luma_reserved", "u1"), ("luma_offsets", "i1", 4),
                      ("chroma_type", "u1"), ("chroma_eo_class", "u1"),
                      ("chroma_band_position", "u1"), ("chroma_reserved", "u1"),
                      ("chroma_offsets", "i1", 4)])
assert SAO_DTYPE.itemsize == 20

# Fields the reference itself exposes (evx_block_desc + selection SAD).
REF_FIELDS = ("block_type", "prediction_mode", "prediction_target", "prediction_target_b",
              "motion_x", "motion_y", "motion_x_b", "motion_y_b", "sp_pred", "sp_amount",
              "sp_index", "sp_pred_b", "sp_amount_b", "sp_index_b", "q_index", "variance",
              "sad", "sad_b", "flags")


class FrameJob(C.Structure):
    _fields_ = [("type", C.c_int32), ("quality", C.c_int32), ("src", Planes),
                ("n_refs", C.c_int32), ("refs", Planes * 4), ("target_ids", C.c_uint8 * 4),
                ("prev", C.c_void_p)]


class EncodeJob(C.Structure):
    _fields_ = [("type", C.c_int32), ("quality", C.c_int32), ("src", Planes),
                ("n_refs", C.c_int32), ("refs", Planes * 2), ("target_ids", C.c_uint8 * 2),
                ("prev", C.c_void_p)]


def _ptr16(a):
    assert a.dtype == np.int16 and a.flags.c_contiguous
    return a.ctypes.data_as(I16P)


class Picture:
    """A 4:2:0 int16 picture; keeps the arrays alive for the Planes struct."""

    def __init__(self, y, u, v):
        self.y = np.ascontiguousarray(y, dtype=np.int16)
        self.u = np.ascontiguousarray(u, dtype=np.int16)
        self.v = np.ascontiguousarray(v, dtype=np.int16)
        h, w = self.y.shape
        assert w % 16 == 0 and h % 16 == 0
        assert self.u.shape == (h // 2, w // 2) and self.v.shape == (h // 2, w // 2)
        self.planes = Planes(_ptr16(self.y), _ptr16(self.u), _ptr16(self.v), w, h, w, w // 2)

    @property
    def width(self):
        return self.y.shape[1]

    @property
    def height(self):
        return self.y.shape[0]

    def block(self, x, y):
        out = np.empty(384, np.int16)
        out[:256] = self.y[y:y + 16, x:x + 16].ravel()
        out[256:320] = self.u[y // 2:y // 2 + 8, x // 2:x // 2 + 8].ravel()
        out[320:] = self.v[y // 2:y // 2 + 8, x // 2:x // 2 + 8].ravel()
        return out


class _Lib:
    prefix = ""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)

    def f(self, name, restype, *argtypes):
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = restype
        fn.argtypes = argtypes
        return fn


class _Recon:
    """Reconstruction entry points shared by the restatement and the
    reference build (same names and argument layout)."""

    def _bind_recon(self):
        self._deq = self.f("dequantize_inter_mb", None, C.c_int32, I16P, I16P)
        self._idct = self.f("inverse_transform_add_mb", None, I16P, I16P, I16P)
        self._recon = self.f("reconstruct_frame", C.c_int32, C.POINTER(FrameJob), C.c_void_p,
                             I16P, I16P, I16P, I16P)
        self._dbl_line = self.f("deblock_line", None, I16P, C.c_int32, C.c_int32, C.c_int32)
        self._dbl_frame = self.f("deblock_frame", None, C.c_void_p, C.c_int32, C.c_int32, I16P,
                                 I16P, I16P)
        self._sao_filter = self.f("sao_filter_frame", None, C.c_void_p, C.c_int32, C.c_int32,
                                  I16P, I16P, I16P)
        self._sao_encode = self.f("sao_encode_frame", None, I16P, I16P, I16P, I16P, I16P, I16P,
                                  C.c_int32, C.c_int32, C.c_float, C.c_void_p)

    def dequantize(self, qp, levels):
        lv = np.ascontiguousarray(levels, np.int16)
        out = np.empty(384, np.int16)
        self._deq(qp, _ptr16(lv), _ptr16(out))
        return out

    def idct_add(self, coeffs, pred):
        c, p = (np.ascontiguousarray(a, np.int16) for a in (coeffs, pred))
        out = np.empty(384, np.int16)
        self._idct(_ptr16(c), _ptr16(p), _ptr16(out))
        return out

    def reconstruct_frame(self, ftype, quality, src, refs, records, levels, target_ids=None):
        """(y, u, v) planes of the frame `records` and `levels` describe
        (src only gives the size)."""
        job = make_job(ftype, quality, src, refs, None, target_ids)
        w, h = src.width, src.height
        y = np.zeros((h, w), np.int16)
        u = np.zeros((h // 2, w // 2), np.int16)
        v = np.zeros((h // 2, w // 2), np.int16)
        rec = np.ascontiguousarray(records)
        lv = np.ascontiguousarray(levels, np.int16)
        rc = self._recon(C.byref(job), rec.ctypes.data, _ptr16(lv), _ptr16(y), _ptr16(u),
                         _ptr16(v))
        if rc != 0:
            raise RuntimeError(f"reconstruct_frame failed: {rc}")
        return y, u, v

    def deblock_line(self, v, qp, strength, is_luma):
        """The 8 samples p3 .. q3 of one line after the edge filter."""
        a = np.array(v, np.int16)
        self._dbl_line(_ptr16(a), qp, strength, int(bool(is_luma)))
        return a

    def sao_filter_frame(self, table, y, u, v):
        """Copies of the planes after SAO per `table` (SAO_DTYPE, one per MB)."""
        h, w = y.shape
        t = np.ascontiguousarray(table, SAO_DTYPE)
        if t.size != (w // 16) * (h // 16):
            raise ValueError("one SAO entry per MB")
        oy, ou, ov = (np.array(a, np.int16, order="C") for a in (y, u, v))
        self._sao_filter(t.ctypes.data, w, h, _ptr16(oy), _ptr16(ou), _ptr16(ov))
        return oy, ou, ov

    def sao_encode_frame(self, src, rec, lam):
        """(table, y, u, v): the SAO decisions of the frame whose source
        planes are `src` and deblocked planes `rec`, and rec after SAO."""
        h, w = rec[0].shape
        s3 = [np.ascontiguousarray(a, np.int16) for a in src]
        oy, ou, ov = (np.array(a, np.int16, order="C") for a in rec)
        t = np.zeros((w // 16) * (h // 16), SAO_DTYPE)
        self._sao_encode(*(_ptr16(a) for a in s3), _ptr16(oy), _ptr16(ou), _ptr16(ov), w, h,
                         C.c_float(lam), t.ctypes.data)
        return t, oy, ou, ov

    def deblock_frame(self, records, y, u, v):
        """Deblocked copies of the planes (y: h x w, u/v: h/2 x w/2); the
        records give each MB's block type and q_index."""
        h, w = y.shape
        rec = np.ascontiguousarray(records)
        if rec.size != (w // 16) * (h // 16):
            raise ValueError("one record per MB")
        oy, ou, ov = (np.array(a, np.int16, order="C") for a in (y, u, v))
        self._dbl_frame(rec.ctypes.data, w, h, _ptr16(oy), _ptr16(ou), _ptr16(ov))
        return oy, ou, ov


class _Encode:
    """Mode decision and in-loop encode entry points shared by the
    restatement and the reference build (same names and layout)."""

    def _bind_encode(self, encode_extra=()):
        P = C.POINTER(Planes)
        self._ipred = self.f("intra_prediction", None, C.c_int32, P, C.c_int32, C.c_int32, I16P)
        self._icand = self.f("intra_mode_candidates", None, P, C.c_int32, C.c_int32, C.c_void_p)
        self._isel = self.f("select_intra_mode", C.c_int32, I16P, P, C.c_int32, C.c_int32, I32P)
        self._quant_mb = self.f("quantize_mb", None, C.c_int32, C.c_int32, I16P, I16P)
        self._dequant_mb = self.f("dequantize_mb", None, C.c_int32, C.c_int32, I16P, I16P)
        self._trellis = self.f("trellis_mb", None, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                               I16P, I16P)
        self._encode = self.f("encode_frame", C.c_int32, C.POINTER(EncodeJob), I16P, I16P, I16P,
                              C.c_void_p, I16P, *encode_extra)

    def intra_prediction(self, mode, pic, x, y):
        out = np.empty(384, np.int16)
        self._ipred(mode, C.byref(pic.planes), x, y, _ptr16(out))
        return out

    def intra_mode_candidates(self, pic, x, y):
        out = np.zeros(3, np.uint8)
        self._icand(C.byref(pic.planes), x, y, out.ctypes.data)
        return [int(v) for v in out]

    def select_intra_mode(self, src_blk, pic, x, y):
        """(mode, luma SAD) of select_best_intra_mode."""
        sad = C.c_int32()
        s = np.ascontiguousarray(src_blk, np.int16)
        mode = self._isel(_ptr16(s), C.byref(pic.planes), x, y, C.byref(sad))
        return mode, sad.value

    def quantize_mb(self, qp, block_type, coeffs):
        out = np.empty(384, np.int16)
        self._quant_mb(qp, block_type, _ptr16(np.ascontiguousarray(coeffs, np.int16)),
                       _ptr16(out))
        return out

    def dequantize_mb(self, qp, block_type, levels):
        out = np.empty(384, np.int16)
        self._dequant_mb(qp, block_type, _ptr16(np.ascontiguousarray(levels, np.int16)),
                         _ptr16(out))
        return out

    def trellis(self, qp, block_type, intra_mode, lam, coeffs, levels):
        out = np.array(levels, np.int16, order="C")
        self._trellis(qp, block_type, intra_mode, C.c_float(lam),
                      _ptr16(np.ascontiguousarray(coeffs, np.int16)), _ptr16(out))
        return out

    def _encode_call(self, job, y, u, v, recs, levels, coeffs=None):
        return self._encode(job, _ptr16(y), _ptr16(u), _ptr16(v), recs.ctypes.data,
                            _ptr16(levels), None if coeffs is None else _ptr16(coeffs))

    def encode_frame_coeffs(self, ftype, quality, src, refs, pic, prev=None, target_ids=None):
        """encode_frame, plus the coefficients each MB's levels quantise
        (n_mbs x 384, zero for copies): (records, levels, planes, coeffs)."""
        n = (src.width // 16) * (src.height // 16)
        co = np.zeros((n, 384), np.int16)
        recs, lv, pic = self.encode_frame(ftype, quality, src, refs, pic, prev, target_ids,
                                          _coeffs=co)
        return recs, lv, pic, co

    def encode_frame(self, ftype, quality, src, refs, pic, prev=None, target_ids=None,
                     _coeffs=None):
        """encode_slice of one frame (0 = I, 1 = P with refs[0], 2 = B with
        refs[0] forward and refs[1] backward).  pic: the (y, u, v) planes the
        picture-buffer slot holds before the frame.  Returns (records,
        levels, (y, u, v) reconstruction)."""
        job = EncodeJob()
        job.type = ftype
        job.quality = quality
        job.src = src.planes
        job.n_refs = len(refs)
        for k, r in enumerate(refs):
            job.refs[k] = r.planes
        ids = target_ids if target_ids is not None else [1, 2][:len(refs)]
        for k, t in enumerate(ids):
            job.target_ids[k] = t
        pv = None if prev is None else np.ascontiguousarray(prev)
        job.prev = None if pv is None else pv.ctypes.data
        y, u, v = (np.array(a, np.int16, order="C") for a in pic)
        n = (src.width // 16) * (src.height // 16)
        recs = np.zeros(n, RECORD_DTYPE)
        levels = np.zeros((n, 384), np.int16)
        if _coeffs is None:
            rc = self._encode_call(C.byref(job), y, u, v, recs, levels)
        else:
            rc = self._encode_call(C.byref(job), y, u, v, recs, levels, _coeffs)
        if rc != 0:
            raise RuntimeError(f"encode_frame failed: {rc}")
        return recs, levels, (y, u, v)


class Oracle(_Lib, _Recon, _Encode):
    """The C restatement."""
    prefix = "evo_"

    def __init__(self, path=ORACLE_SO):
        super().__init__(path)
        P = C.POINTER(Planes)
        self._lambda = self.f("lambda", C.c_float, C.c_int32)
        self._interp = self.f("interp_block", None, P, C.c_int32, C.c_int32, C.c_int32,
                              C.c_int32, C.c_int32, I16P)
        self._had = self.f("hadamard_4x4", None, I16P, C.c_int32, I32P)
        self._sad = self.f("block_sad", C.c_int32, I16P, I16P)
        self._mad = self.f("block_mad", C.c_int32, I16P, I16P)
        self._satd = self.f("block_satd", C.c_int32, I16P, I16P)
        self._pred = self.f("mv_predictor", None, C.c_void_p, C.c_uint32, C.c_uint32,
                            C.c_void_p, I16P, I16P)
        self._search = self.f("inter_search", C.c_int32, I16P, P, C.c_int32, C.c_int32,
                              C.c_int32, C.c_int16, C.c_int16, C.POINTER(SearchResult))
        self._intra = self.f("intra_search", C.c_int32, I16P, P, C.c_int32, C.c_int32, C.c_int32,
                             C.POINTER(SearchResult))
        self._predict = self.f("predict_block", None, P, C.c_int32, C.c_int32,
                               C.POINTER(SearchResult), I16P)
        self._bidir = self.f("bidir_average", C.c_int32, I16P, I16P, I16P, I16P)
        self._subt = self.f("sub_transform_mb", None, I16P, I16P, I16P)
        self._var2 = self.f("variance2", C.c_int32, I16P)
        self._qp = self.f("block_qp", C.c_int32, C.c_int32, I16P)
        self._quant = self.f("quantize_inter_mb", None, C.c_int32, I16P, I16P)
        self._frame = self.f("analyze_frame", C.c_int32, C.POINTER(FrameJob), C.c_void_p,
                             I16P, I16P)
        self._bind_recon()
        self._bind_encode((I16P,))

    def lam(self, q):
        return self._lambda(q)

    def interp(self, pic, bx, by, dx, dy, quarter):
        out = np.empty(384, np.int16)
        self._interp(C.byref(pic.planes), bx, by, dx, dy, int(quarter), _ptr16(out))
        return out

    def hadamard(self, blk4x4):
        b = np.ascontiguousarray(blk4x4, np.int16).ravel()
        out = np.empty(16, np.int32)
        self._had(_ptr16(b), 4, out.ctypes.data_as(I32P))
        return out

    def metrics(self, a, b):
        a = np.ascontiguousarray(a, np.int16)
        b = np.ascontiguousarray(b, np.int16)
        return (self._sad(_ptr16(a), _ptr16(b)), self._mad(_ptr16(a), _ptr16(b)),
                self._satd(_ptr16(a), _ptr16(b)))

    def mv_predictor(self, table, wb, idx, prev=None):
        px, py = C.c_int16(), C.c_int16()
        self._pred(table.ctypes.data, wb, idx, None if prev is None else prev.ctypes.data,
                   C.byref(px), C.byref(py))
        return px.value, py.value

    def inter_search(self, src_blk, ref, x, y, quality, pmx, pmy):
        r = SearchResult()
        s = np.ascontiguousarray(src_blk, np.int16)
        self._search(_ptr16(s), C.byref(ref.planes), x, y, quality, pmx, pmy, C.byref(r))
        return r

    def intra_search(self, src_blk, pic, x, y, quality):
        """calculate_intra_prediction of the MB at (x, y) over `pic`, the
        picture being reconstructed."""
        r = SearchResult()
        s = np.ascontiguousarray(src_blk, np.int16)
        self._intra(_ptr16(s), C.byref(pic.planes), x, y, quality, C.byref(r))
        return r

    def predict(self, ref, x, y, r):
        out = np.empty(384, np.int16)
        self._predict(C.byref(ref.planes), x, y, C.byref(r), _ptr16(out))
        return out

    def bidir(self, src, f, b):
        avg = np.empty(384, np.int16)
        src, f, b = (np.ascontiguousarray(a, np.int16) for a in (src, f, b))
        sad = self._bidir(_ptr16(src), _ptr16(f), _ptr16(b), _ptr16(avg))
        return sad, avg

    def sub_transform(self, src, pred):
        out = np.empty(384, np.int16)
        src, pred = (np.ascontiguousarray(a, np.int16) for a in (src, pred))
        self._subt(_ptr16(src), _ptr16(pred), _ptr16(out))
        return out

    def variance2(self, coeffs):
        c = np.ascontiguousarray(coeffs, np.int16)
        return self._var2(_ptr16(c))

    def block_qp(self, quality, coeffs):
        c = np.ascontiguousarray(coeffs, np.int16)
        return self._qp(quality, _ptr16(c))

    def quantize(self, qp, coeffs):
        c = np.ascontiguousarray(coeffs, np.int16)
        out = np.empty(384, np.int16)
        self._quant(qp, _ptr16(c), _ptr16(out))
        return out

    def analyze_frame(self, ftype, quality, src, refs, prev=None, target_ids=None):
        return _run_frame(self._frame, ftype, quality, src, refs, prev, target_ids)


def make_job(ftype, quality, src, refs, prev, target_ids):
    job = FrameJob()
    job.type = ftype
    job.quality = quality
    job.src = src.planes
    job.n_refs = len(refs)
    for k, r in enumerate(refs):
        job.refs[k] = r.planes
    ids = target_ids if target_ids is not None else (
        list(range(1, len(refs) + 1)) if ftype == 1 else [0, 1])
    for k, t in enumerate(ids):
        job.target_ids[k] = t
    job.prev = None if prev is None else prev.ctypes.data
    return job


def _run_frame(fn, ftype, quality, src, refs, prev, target_ids):
    n = (src.width // 16) * (src.height // 16)
    recs = np.zeros(n, RECORD_DTYPE)
    coeffs = np.zeros((n, 384), np.int16)
    levels = np.zeros((n, 384), np.int16)
    job = make_job(ftype, quality, src, refs, prev, target_ids)
    rc = fn(C.byref(job), recs.ctypes.data, _ptr16(coeffs), _ptr16(levels))
    if rc != 0:
        raise RuntimeError(f"analyze_frame failed: {rc}")
    return recs, coeffs, levels


class Reference(_Lib, _Recon, _Encode):
    """The reference's own functions (oracle/_ref, built from its sources)."""
    prefix = "evr_"

    def __init__(self, path=REF_SO):
        super().__init__(path)
        P = C.POINTER(Planes)
        self._lambda = self.f("lambda", C.c_float, C.c_int32)
        self._interp_many = self.f("interp_many", None, P, C.c_int32, I32P, I32P, I32P, I32P,
                                   I32P, I16P)
        self._had = self.f("hadamard_4x4", None, I16P, C.c_int32, I32P)
        self._metrics = self.f("block_metrics", None, I16P, I16P, I32P, I32P, I32P)
        self._subt = self.f("sub_transform_mb", None, I16P, I16P, I16P)
        self._var2 = self.f("variance2", C.c_int32, I16P)
        self._qp = self.f("block_qp", C.c_int32, C.c_int32, I16P)
        self._quant = self.f("quantize_mb", None, C.c_int32, C.c_int32, I16P, I16P)
        self._pred = self.f("mv_predictor", None, C.c_void_p, C.c_uint32, C.c_uint32,
                            C.c_uint32, C.c_void_p, I16P, I16P)
        self._search = self.f("inter_search", C.c_int32, I16P, P, C.c_int32, C.c_int32,
                              C.c_int32, C.c_int16, C.c_int16, C.c_void_p)
        self._intra = self.f("intra_search", C.c_int32, I16P, P, C.c_int32, C.c_int32, C.c_int32,
                             C.c_void_p)
        self._bidir = self.f("bidir", C.c_int32, I16P, P, P, C.c_int32, C.c_int32, C.c_int32,
                             C.c_int16, C.c_int16, C.c_void_p, I16P)
        self._frame = self.f("analyze_frame", C.c_int32, C.POINTER(FrameJob), C.c_void_p,
                             I16P, I16P)
        self._frame_wpp = self.f("analyze_frame_wpp", C.c_int32, C.POINTER(FrameJob),
                                 C.c_void_p, I16P, I16P, C.c_int32)
        self._bind_recon()
        self._bind_encode((C.c_int32,))
        self.encode_threads = 1

    def _encode_call(self, job, y, u, v, recs, levels):
        return self._encode(job, _ptr16(y), _ptr16(u), _ptr16(v), recs.ctypes.data,
                            _ptr16(levels), self.encode_threads)

    def lam(self, q):
        return self._lambda(q)

    def analyze_frame_serialized(self, ftype, quality, src, refs, prev=None, target_ids=None):
        """The unified block table (serialize.cpp:184-314) the reference's
        own analysis leaves, arithmetic-coded as serialize_slice writes it."""
        if not hasattr(self, "_afs"):
            self._afs = self.f("analyze_frame_serialized", C.c_int32, C.POINTER(FrameJob),
                               C.c_void_p, C.c_int32)
        job = make_job(ftype, quality, src, refs, prev, target_ids)
        out = np.zeros(1 << 22, np.uint8)
        n = self._afs(C.byref(job), out.ctypes.data, out.size)
        if n < 0:
            raise RuntimeError(f"analyze_frame_serialized failed: {n}")
        return out[:n].tobytes()

    def serialize_records(self, records, width, height, ftype, prev=None, target_ids=None):
        """The same serialization from block records (target ids mapped to
        reference indexes as make_job assigns them)."""
        if not hasattr(self, "_ser"):
            self._ser = self.f("serialize_records", C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                               C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                               C.c_int32)
        n_refs = 2 if ftype == 2 else 4
        ids = target_ids if target_ids is not None else (
            list(range(1, n_refs + 1)) if ftype == 1 else [0, 1])
        tid = np.array(ids, np.uint8)
        rec = np.ascontiguousarray(records)
        pv = None if prev is None else np.ascontiguousarray(prev)
        out = np.zeros(1 << 22, np.uint8)
        n = self._ser(rec.ctypes.data, width // 16, height // 16,
                      None if pv is None else pv.ctypes.data, int(ftype == 2), tid.ctypes.data,
                      len(ids), out.ctypes.data, out.size)
        if n < 0:
            raise RuntimeError("serialize_records failed")
        return out[:n].tobytes()

    def serialize_sao(self, table, width, height):
        """serialize_sao_table (serialize.cpp:564-622) of an SAO table."""
        if not hasattr(self, "_sers"):
            self._sers = self.f("serialize_sao", C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                C.c_void_p, C.c_int32)
        t = np.ascontiguousarray(table, SAO_DTYPE)
        out = np.zeros(1 << 20, np.uint8)
        n = self._sers(t.ctypes.data, width // 16, height // 16, out.ctypes.data, out.size)
        if n < 0:
            raise RuntimeError("serialize_sao failed")
        return out[:n].tobytes()

    def analyze_frame_wpp(self, ftype, quality, src, refs, prev=None, target_ids=None,
                          threads=1):
        """analyze_frame on `threads` threads in the reference's wavefront row
        order (identical results)."""
        def fn(job, rec, co, lv):
            return self._frame_wpp(job, rec, co, lv, threads)
        return _run_frame(fn, ftype, quality, src, refs, prev, target_ids)

    def interp_many(self, pic, bx, by, dx, dy, quarter):
        arrs = [np.ascontiguousarray(a, np.int32) for a in (bx, by, dx, dy, quarter)]
        n = len(arrs[0])
        out = np.empty((n, 384), np.int16)
        self._interp_many(C.byref(pic.planes), n, *[a.ctypes.data_as(I32P) for a in arrs],
                          _ptr16(out))
        return out

    def hadamard(self, blk4x4):
        b = np.ascontiguousarray(blk4x4, np.int16).ravel()
        out = np.empty(16, np.int32)
        self._had(_ptr16(b), 4, out.ctypes.data_as(I32P))
        return out

    def metrics(self, a, b):
        a = np.ascontiguousarray(a, np.int16)
        b = np.ascontiguousarray(b, np.int16)
        s, m, t = C.c_int32(), C.c_int32(), C.c_int32()
        self._metrics(_ptr16(a), _ptr16(b), C.byref(s), C.byref(m), C.byref(t))
        return s.value, m.value, t.value

    def sub_transform(self, src, pred):
        out = np.empty(384, np.int16)
        src, pred = (np.ascontiguousarray(a, np.int16) for a in (src, pred))
        self._subt(_ptr16(src), _ptr16(pred), _ptr16(out))
        return out

    def variance2(self, coeffs):
        return self._var2(_ptr16(np.ascontiguousarray(coeffs, np.int16)))

    def block_qp(self, quality, coeffs):
        return self._qp(quality, _ptr16(np.ascontiguousarray(coeffs, np.int16)))

    def quantize(self, qp, coeffs, block_type=2):
        out = np.empty(384, np.int16)
        self._quant(qp, block_type, _ptr16(np.ascontiguousarray(coeffs, np.int16)), _ptr16(out))
        return out

    def mv_predictor(self, table, wb, idx, prev=None):
        px, py = C.c_int16(), C.c_int16()
        self._pred(table.ctypes.data, wb, len(table), idx,
                   None if prev is None else prev.ctypes.data, C.byref(px), C.byref(py))
        return px.value, py.value

    def inter_search(self, src_blk, ref, x, y, quality, pmx, pmy):
        rec = np.zeros(1, RECORD_DTYPE)
        s = np.ascontiguousarray(src_blk, np.int16)
        self._search(_ptr16(s), C.byref(ref.planes), x, y, quality, pmx, pmy, rec.ctypes.data)
        return rec[0]

    def intra_search(self, src_blk, pic, x, y, quality):
        rec = np.zeros(1, RECORD_DTYPE)
        s = np.ascontiguousarray(src_blk, np.int16)
        self._intra(_ptr16(s), C.byref(pic.planes), x, y, quality, rec.ctypes.data)
        return rec[0]

    def bidir(self, src_blk, fwd, bwd, x, y, quality, pmx, pmy):
        rec = np.zeros(1, RECORD_DTYPE)
        avg = np.empty(384, np.int16)
        s = np.ascontiguousarray(src_blk, np.int16)
        self._bidir(_ptr16(s), C.byref(fwd.planes), C.byref(bwd.planes), x, y, quality, pmx, pmy,
                    rec.ctypes.data, _ptr16(avg))
        return rec[0], avg

    def analyze_frame(self, ftype, quality, src, refs, prev=None, target_ids=None):
        return _run_frame(self._frame, ftype, quality, src, refs, prev, target_ids)


def _words_table(words):
    """A record table holding only block types (motion bit) and MVs."""
    t = np.zeros(len(words), RECORD_DTYPE)
    for i, (m, mx, my) in enumerate(words):
        t["block_type"][i] = 2 if m else 0
        t["motion_x"][i] = mx
        t["motion_y"][i] = my
    return t


def sweep_model(o, ftype, quality, src, refs, prev=None):
    """Model of the GPU's sweep mode on the oracle's searches: Jacobi sweeps
    to the fixpoint of "each MB's record is its searches under the predictor
    of its neighbours' records" (compute_mv_predictor, motion.cpp:271-346,
    read by perform_inter_search only through motion.cpp:613-635).

    Sweep 0 searches every MB with the temporal table (P frames) standing in
    for the neighbours (the GPU's deterministic setting, sweep_reads 1);
    sweep s reads every MB's word (motion bit, MV) as of sweep s - 1,
    searches a listed MB again only if its predictor key (absent, or the
    clamped candidate position) changed, and a MB whose word changed lists
    itself and its right, bottom-left, bottom and bottom-right neighbours for
    sweep s + 1; sweep 1 lists every MB.  Selection follows encode.cpp:183-248 (P,
    against an empty sentinel) and 386-443 (B).

    Returns (final words per MB as (motion, mv_x, mv_y), MBs listed per sweep
    (sweep 0: all), MBs searched per sweep); the words must equal the
    sequential records'."""
    wb, hb = src.width // 16, src.height // 16
    n = wb * hb
    W, H = src.width, src.height
    lam = np.float32(o.lam(quality))
    tprev = prev if ftype == 1 else None

    def word_of_rec(r):
        return (1, int(r["motion_x"]), int(r["motion_y"])) if r["block_type"] & 2 else (0, 0, 0)

    guess = [word_of_rec(r) for r in tprev] if tprev is not None else [(0, 0, 0)] * n

    def decide(m, pmx, pmy):
        x, y = (m % wb) * 16, (m // wb) * 16
        blk = src.block(x, y)
        res = [o.inter_search(blk, r, x, y, quality, pmx, pmy) for r in refs]
        if ftype == 1:
            best, best_sad, best_copy = -1, 2**31 - 1, False
            for k, r in enumerate(res):
                bias = int(lam * np.float32((k + 1) * 2))
                copy = (r.block_type & 4) != 0
                if copy != best_copy:
                    if copy:
                        best, best_sad, best_copy = k, r.sad, True
                elif r.sad + bias < best_sad:
                    best, best_sad, best_copy = k, r.sad, copy
            r = res[best]
            return (1, r.motion_x, r.motion_y) if r.block_type & 2 else (0, 0, 0)
        f, b = res
        best, best_sad = -1, 2**31 - 1
        if f.sad < best_sad:
            best, best_sad = 0, f.sad
        if b.sad < best_sad:
            best, best_sad = 1, b.sad
        bsad, _ = o.bidir(blk, o.predict(refs[0], x, y, f), o.predict(refs[1], x, y, b))
        if bsad + int(lam * np.float32(8.0)) < best_sad:
            return (1, f.motion_x, f.motion_y)
        r = res[best]
        return (1, r.motion_x, r.motion_y) if r.block_type & 2 else (0, 0, 0)

    def key_of(m, pmx, pmy):
        if pmx == 0 and pmy == 0:
            return None
        x, y = (m % wb) * 16, (m // wb) * 16
        return (min(max(x + pmx, 0), W - 16), min(max(y + pmy, 0), H - 16))

    def dependants(m):
        bx, by = m % wb, m // wb
        out = [m]
        if bx + 1 < wb:
            out.append(m + 1)
        if by + 1 < hb:
            if bx > 0:
                out.append(m + wb - 1)
            out.append(m + wb)
            if bx + 1 < wb:
                out.append(m + wb + 1)
        return out

    words = [list(guess), list(guess)]
    keys = [None] * n
    n_listed, searched = [], []
    tab = _words_table(guess)
    for m in range(n):
        pmx, pmy = o.mv_predictor(tab, wb, m, tprev)
        keys[m] = key_of(m, pmx, pmy)
        words[0][m] = decide(m, pmx, pmy)
    n_listed.append(n)
    searched.append(n)
    listed = set(range(n))      # sweep 1 checks every MB
    s = 1
    while listed:
        old = words[(s - 1) & 1]
        cur = words[s & 1]
        tab = _words_table(old)
        nxt = set()
        cnt = 0
        for m in sorted(listed):
            pmx, pmy = o.mv_predictor(tab, wb, m, tprev)
            k = key_of(m, pmx, pmy)
            if k == keys[m]:
                cur[m] = old[m]
                continue
            keys[m] = k
            cnt += 1
            w = decide(m, pmx, pmy)
            cur[m] = w
            if w != old[m]:
                nxt.update(dependants(m))
        n_listed.append(len(listed))
        searched.append(cnt)
        listed = nxt
        s += 1
    return words[(s - 1) & 1], n_listed, searched


def reference_available():
    return os.path.exists(REF_SO)
