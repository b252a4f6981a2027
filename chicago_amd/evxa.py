"""Python host binding of libevxa.so (the C ABI in include/evxa.h).

Mirrors the reference encoder's inter-analysis interface: a `Context` owns a
device queue and an HBM budget, a `Stream` owns device-resident picture slots
(the reference's prediction_cache, common.h:124-137) and record tables, and
frame jobs replace classify_block / classify_block_bidir / encode_block's inter
analysis (encode.cpp:168-739).  Per-block entry points mirror
interpolate_subpixel_6tap (interp.cpp:166), perform_inter_search
(motion.cpp:567) and the sub_transform / QP / quantize chain
(macroblock.h:272, quantize.cpp:97, 454).

There is no CPU fallback: importing works anywhere, but every call needs the
built library and a GPU, and fails loudly otherwise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libevxa.so")

FRAME_I = 0
FRAME_P = 1
FRAME_B = 2
MB_SAMPLES = 384

I16P = C.POINTER(C.c_int16)
I32P = C.POINTER(C.c_int32)

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

# evxa_sao_info: one per MB (evxa.h; sao.h:62-74).
SAO_DTYPE = np.dtype([("merge_left", "u1"), ("merge_above", "u1"), ("reserved", "u1", 2),
                      ("luma_type", "u1"), ("luma_eo_class", "u1"), ("luma_band_position", "u1"),
                      ("luma_reserved", "u1"), ("luma_offsets", "i1", 4),
                      ("chroma_type", "u1"), ("chroma_eo_class", "u1"),
                      ("chroma_band_position", "u1"), ("chroma_reserved", "u1"),
                      ("chroma_offsets", "i1", 4)])
assert SAO_DTYPE.itemsize == 20

SEARCH_DTYPE = np.dtype([(n, "<i4") for n in (
    "best_x", "best_y", "sad", "mad", "ssd")] + [("cost_bits", "<u4")] + [(n, "<i4") for n in (
        "sp_enabled", "sp_amount", "sp_index", "block_type", "motion_x", "motion_y", "n_int",
        "n_axial", "n_diag")])
assert SEARCH_DTYPE.itemsize == 60


COMPLETION_DTYPE = np.dtype([("tag", "<u8"), ("status", "<i4"), ("batch_frames", "<i4"),
                             ("t_arrival", "<f8"), ("t_launch", "<f8"), ("t_done", "<f8"),
                             ("device_ms", "<f4"), ("reserved", "<i4")])
assert COMPLETION_DTYPE.itemsize == 48


MODE_WAVEFRONT = 1
MODE_SWEEP = 2


SWEEP_READS_GUESS = 1
SWEEP_READS_FRESH = 2
SWEEP_ORDER_RASTER = 1
SWEEP_ORDER_DIAGONAL = 2


class Tuning(C.Structure):
    _fields_ = [("waves_per_search", C.c_int32), ("max_groups", C.c_int32), ("lists", C.c_int32),
                ("cap_pct", C.c_int32), ("mode", C.c_int32), ("sweep_reads", C.c_int32),
                ("sweep_wait_us", C.c_int32), ("sweep_order", C.c_int32),
                ("claim_batch", C.c_int32), ("window_reach", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class FrameDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("quality", C.c_int32), ("src_slot", C.c_int32),
                ("n_refs", C.c_int32), ("ref_slots", C.c_int32 * 4),
                ("target_ids", C.c_uint8 * 4), ("temporal_mvp", C.c_int32),
                ("update_temporal", C.c_int32)]


class EncodeDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("quality", C.c_int32), ("src_slot", C.c_int32),
                ("dst_slot", C.c_int32), ("ref_slots", C.c_int32 * 2),
                ("target_ids", C.c_uint8 * 2), ("reserved", C.c_uint8 * 2),
                ("temporal_mvp", C.c_int32), ("update_temporal", C.c_int32)]


# Every symbol include/evxa.h declares.
EXPORTS = (
    "evxa_create", "evxa_destroy", "evxa_last_error", "evxa_version", "evxa_stream_bytes",
    "evxa_context_bytes", "evxa_lambda", "evxa_stream_open", "evxa_stream_close",
    "evxa_upload_picture", "evxa_stream_reset_temporal", "evxa_analyze_frame",
    "evxa_analyze_batch", "evxa_synchronize", "evxa_last_launch_ms", "evxa_launch_times",
    "evxa_read_results", "evxa_download_picture", "evxa_reconstruct", "evxa_deblock",
    "evxa_reconstruct_deblocked", "evxa_sao", "evxa_sao_filter",
    "evxa_graph_create", "evxa_graph_launch", "evxa_graph_destroy", "evxa_interp_blocks",
    "evxa_search_blocks", "evxa_dct_quant_blocks", "evxa_clock", "evxa_server_create",
    "evxa_server_submit", "evxa_server_poll", "evxa_server_pending", "evxa_server_destroy",
    "evxa_set_tuning", "evxa_get_tuning", "evxa_sweep_stats", "evxa_analyze_frame_noalloc",
    "evxa_stream_wait", "evxa_stream_query", "evxa_intra_search_blocks", "evxa_encode_frame",
    "evxa_encode_batch")


class EvxaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"evxa error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libevxa.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m chicago_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = {
        "evxa_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(vp)]),
        "evxa_destroy": (C.c_int, [vp]),
        "evxa_last_error": (C.c_char_p, []),
        "evxa_version": (C.c_int32, []),
        "evxa_stream_bytes": (C.c_uint64, [C.c_int32, C.c_int32, C.c_int32]),
        "evxa_context_bytes": (C.c_uint64, [vp]),
        "evxa_lambda": (C.c_float, [C.c_int32]),
        "evxa_stream_open": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)]),
        "evxa_stream_close": (C.c_int, [vp]),
        "evxa_upload_picture": (C.c_int, [vp, C.c_int32, I16P, I16P, I16P, C.c_int32, C.c_int32]),
        "evxa_stream_reset_temporal": (C.c_int, [vp]),
        "evxa_analyze_frame": (C.c_int, [vp, C.POINTER(FrameDesc)]),
        "evxa_analyze_frame_noalloc": (C.c_int, [vp, C.POINTER(FrameDesc)]),
        "evxa_encode_frame": (C.c_int, [vp, C.POINTER(EncodeDesc)]),
        "evxa_encode_batch": (C.c_int, [vp, C.POINTER(vp), C.POINTER(EncodeDesc), C.c_int32]),
        "evxa_analyze_batch": (C.c_int, [vp, C.POINTER(vp), C.POINTER(FrameDesc), C.c_int32]),
        "evxa_synchronize": (C.c_int, [vp]),
        "evxa_stream_wait": (C.c_int, [vp]),
        "evxa_stream_query": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "evxa_last_launch_ms": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "evxa_launch_times": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int32,
                                        C.POINTER(C.c_int32)]),
        "evxa_read_results": (C.c_int, [vp, vp, I16P, I16P]),
        "evxa_download_picture": (C.c_int, [vp, C.c_int32, I16P, I16P, I16P, C.c_int32, C.c_int32]),
        "evxa_reconstruct": (C.c_int, [vp, C.c_int32, I16P]),
        "evxa_deblock": (C.c_int, [vp, C.c_int32, vp]),
        "evxa_reconstruct_deblocked": (C.c_int, [vp, C.c_int32, I16P]),
        "evxa_sao": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, vp]),
        "evxa_sao_filter": (C.c_int, [vp, C.c_int32, vp]),
        "evxa_graph_create": (C.c_int, [vp, C.POINTER(vp), C.POINTER(FrameDesc), C.c_int32,
                                        C.POINTER(C.c_int32)]),
        "evxa_graph_launch": (C.c_int, [vp, C.c_int32]),
        "evxa_graph_destroy": (C.c_int, [vp, C.c_int32]),
        "evxa_interp_blocks": (C.c_int, [vp, C.c_int32, C.c_int32, I32P, I32P, I32P, I32P, I32P,
                                         I16P]),
        "evxa_search_blocks": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, I32P,
                                         I32P, I16P, I16P, vp]),
        "evxa_intra_search_blocks": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                               I32P, I32P, vp]),
        "evxa_dct_quant_blocks": (C.c_int, [vp, C.c_int32, I16P, I16P, I32P, I16P, I16P, I32P,
                                            I32P]),
        "evxa_clock": (C.c_double, []),
        "evxa_server_create": (C.c_int, [vp, C.c_int32, C.c_int32, C.POINTER(vp)]),
        "evxa_server_submit": (C.c_int, [vp, vp, C.POINTER(FrameDesc), C.c_uint64, C.c_double]),
        "evxa_server_poll": (C.c_int, [vp, vp, C.c_int32, C.POINTER(C.c_int32)]),
        "evxa_server_pending": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
        "evxa_server_destroy": (C.c_int, [vp]),
        "evxa_set_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "evxa_get_tuning": (C.c_int, [vp, C.POINTER(Tuning), C.POINTER(C.c_int32)]),
        "evxa_sweep_stats": (C.c_int, [vp, I32P, I32P, C.POINTER(C.c_float), C.c_int32,
                                       C.POINTER(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise EvxaError(rc, _lib.evxa_last_error().decode())


def _p16(a):
    assert a.dtype == np.int16 and a.flags.c_contiguous
    return a.ctypes.data_as(I16P)


def _p32(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(I32P)


def lambda_of(quality: int) -> float:
    return load_library().evxa_lambda(quality)


def stream_bytes(width: int, height: int, n_slots: int) -> int:
    return load_library().evxa_stream_bytes(width, height, n_slots)


def frame_desc(ftype, quality, src_slot, ref_slots, target_ids=None, temporal_mvp=True,
               update_temporal=None):
    d = FrameDesc()
    d.type = ftype
    d.quality = quality
    d.src_slot = src_slot
    d.n_refs = len(ref_slots)
    for k, s in enumerate(ref_slots):
        d.ref_slots[k] = s
    ids = target_ids if target_ids is not None else (
        list(range(1, len(ref_slots) + 1)) if ftype == FRAME_P else [0, 1])
    for k, t in enumerate(ids):
        d.target_ids[k] = t
    d.temporal_mvp = int(bool(temporal_mvp)) if ftype == FRAME_P else 0
    if update_temporal is None:
        update_temporal = ftype == FRAME_P
    d.update_temporal = int(bool(update_temporal))
    return d


def encode_desc(ftype, quality, src_slot, dst_slot, ref_slots=(), target_ids=None,
                temporal_mvp=True, update_temporal=None):
    """One frame of the in-loop encode: FRAME_I, FRAME_P (ref_slots[0]) or
    FRAME_B (forward, backward)."""
    d = EncodeDesc()
    d.type = ftype
    d.quality = quality
    d.src_slot = src_slot
    d.dst_slot = dst_slot
    for k, s in enumerate(ref_slots):
        d.ref_slots[k] = s
    ids = target_ids if target_ids is not None else [1, 2][:len(ref_slots)]
    for k, t in enumerate(ids):
        d.target_ids[k] = t
    d.temporal_mvp = int(bool(temporal_mvp)) if ftype == FRAME_P else 0
    if update_temporal is None:
        update_temporal = ftype != FRAME_B
    d.update_temporal = int(bool(update_temporal))
    return d


class Context:
    """A device queue with an HBM admission budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.evxa_create(device, hbm_budget, C.byref(h)))
        self.h = h
        self.streams = []

    def close(self):
        if self.h:
            _check(_lib.evxa_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def bytes_used(self) -> int:
        return _lib.evxa_context_bytes(self.h)

    def set_tuning(self, waves_per_search=0, max_groups=0, lists=0, cap_pct=0, mode=0,
                   sweep_reads=0, sweep_wait_us=-1, sweep_order=0, claim_batch=0,
                   window_reach=0):
        t = Tuning(waves_per_search, max_groups, lists, cap_pct, mode, sweep_reads, sweep_wait_us,
                   sweep_order, claim_batch, window_reach)
        _check(_lib.evxa_set_tuning(self.h, C.byref(t)))

    def sweep_stats(self, max_n=4096, times=False):
        """Per sweep of the last sweep-mode launch: (MBs listed, MBs searched)
        lists, plus the microseconds from the first ticket to each sweep's
        end if `times`; empty after a wavefront launch."""
        t = np.zeros(max_n, np.int32)
        m = np.zeros(max_n, np.int32)
        us = np.zeros(max_n, np.float32)
        n = C.c_int32()
        _check(_lib.evxa_sweep_stats(self.h, _p32(t), _p32(m),
                                     us.ctypes.data_as(C.POINTER(C.c_float)), max_n, C.byref(n)))
        k = min(n.value, max_n)
        if times:
            return t[:k].tolist(), m[:k].tolist(), [round(float(x), 1) for x in us[:k]]
        return t[:k].tolist(), m[:k].tolist()

    def tuning(self):
        """(current Tuning as a dict, workgroups of that shape resident at once)."""
        t = Tuning()
        n = C.c_int32()
        _check(_lib.evxa_get_tuning(self.h, C.byref(t), C.byref(n)))
        return t.as_dict(), n.value

    def open_stream(self, width, height, n_slots):
        return Stream(self, width, height, n_slots)

    def synchronize(self):
        _check(_lib.evxa_synchronize(self.h))

    def last_launch_ms(self) -> float:
        ms = C.c_float()
        _check(_lib.evxa_last_launch_ms(self.h, C.byref(ms)))
        return ms.value

    def launch_times(self, max_n=8192):
        """Device ms of each analysis launch not reported before (oldest first)."""
        buf = (C.c_float * max_n)()
        n = C.c_int32()
        _check(_lib.evxa_launch_times(self.h, buf, max_n, C.byref(n)))
        return np.array(buf[:n.value], np.float64)

    def analyze_batch(self, streams, descs):
        n = len(streams)
        arr = (C.c_void_p * n)(*[s.h for s in streams])
        d = (FrameDesc * n)(*descs)
        _check(_lib.evxa_analyze_batch(self.h, arr, d, n))
        for s, dd in zip(streams, descs):
            s._note(dd)

    def encode_batch(self, streams, descs):
        """evxa_encode_batch: the in-loop encode of one frame of each stream
        in one launch."""
        n = len(streams)
        hs = (C.c_void_p * n)(*[s.h for s in streams])
        ds = (EncodeDesc * n)(*descs)
        _check(_lib.evxa_encode_batch(self.h, hs, ds, n))

    def graph_create(self, streams, descs) -> int:
        n = len(streams)
        arr = (C.c_void_p * n)(*[s.h for s in streams])
        d = (FrameDesc * n)(*descs)
        gid = C.c_int32()
        _check(_lib.evxa_graph_create(self.h, arr, d, n, C.byref(gid)))
        return gid.value

    def graph_launch(self, gid):
        _check(_lib.evxa_graph_launch(self.h, gid))

    def graph_destroy(self, gid):
        _check(_lib.evxa_graph_destroy(self.h, gid))

    def dct_quant_blocks(self, src, pred, quality):
        """sub_transform + QP + variance2 + inter quantisation of n blocks."""
        src = np.ascontiguousarray(src, np.int16).reshape(-1, MB_SAMPLES)
        pred = np.ascontiguousarray(pred, np.int16).reshape(-1, MB_SAMPLES)
        n = src.shape[0]
        q = np.ascontiguousarray(np.broadcast_to(np.asarray(quality, np.int32), (n,)))
        co = np.empty_like(src)
        lv = np.empty_like(src)
        qp = np.empty(n, np.int32)
        var = np.empty(n, np.int32)
        _check(_lib.evxa_dct_quant_blocks(self.h, n, _p16(src), _p16(pred), _p32(q), _p16(co),
                                          _p16(lv), _p32(qp), _p32(var)))
        return co, lv, qp, var


class Stream:
    """Device-resident picture slots and analysis tables of one sequence."""

    def __init__(self, ctx: Context, width: int, height: int, n_slots: int):
        self.ctx = ctx
        self.width, self.height, self.n_slots = width, height, n_slots
        self.n_mbs = (width // 16) * (height // 16)
        h = C.c_void_p()
        _check(_lib.evxa_stream_open(ctx.h, width, height, n_slots, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _check(_lib.evxa_stream_close(self.h))
            self.h = None

    def upload(self, slot, y, u, v):
        y = np.ascontiguousarray(y, np.int16)
        u = np.ascontiguousarray(u, np.int16)
        v = np.ascontiguousarray(v, np.int16)
        assert y.shape == (self.height, self.width)
        assert u.shape == v.shape == (self.height // 2, self.width // 2)
        _check(_lib.evxa_upload_picture(self.h, slot, _p16(y), _p16(u), _p16(v), self.width,
                                        self.width // 2))

    def download(self, slot):
        """(y, u, v) planes of picture slot `slot`."""
        y = np.empty((self.height, self.width), np.int16)
        u = np.empty((self.height // 2, self.width // 2), np.int16)
        v = np.empty((self.height // 2, self.width // 2), np.int16)
        _check(_lib.evxa_download_picture(self.h, slot, _p16(y), _p16(u), _p16(v), self.width,
                                          self.width // 2))
        return y, u, v

    def reconstruct(self, dst_slot, levels=None):
        """Rebuild the last analysed frame into `dst_slot` (prediction plus
        the inverse-quantised, inverse-transformed `levels`, n_mbs x 384 in
        MB order; None: the stream's own levels)."""
        if levels is not None:
            levels = np.ascontiguousarray(levels, np.int16)
            assert levels.shape == (self.n_mbs, MB_SAMPLES)
        _check(_lib.evxa_reconstruct(self.h, dst_slot, None if levels is None else _p16(levels)))

    def reconstruct_deblocked(self, dst_slot, levels=None):
        """reconstruct() then deblock() with the frame's own records, in one
        call (no intermediate copy of the picture)."""
        if levels is not None:
            levels = np.ascontiguousarray(levels, np.int16)
            assert levels.shape == (self.n_mbs, MB_SAMPLES)
        _check(_lib.evxa_reconstruct_deblocked(self.h, dst_slot,
                                               None if levels is None else _p16(levels)))

    def sao(self, slot, src_slot, quality, table=True):
        """The encoder's SAO of the deblocked frame in `slot` against its
        source in `src_slot`, applied in place; returns the decisions
        (SAO_DTYPE, one per MB) unless table is False."""
        t = np.zeros(self.n_mbs, SAO_DTYPE) if table else None
        _check(_lib.evxa_sao(self.h, slot, src_slot, quality, None if t is None else t.ctypes.data))
        return t

    def sao_filter(self, slot, table):
        """Apply an SAO table (SAO_DTYPE, one per MB) to `slot` in place."""
        t = np.ascontiguousarray(table, SAO_DTYPE)
        assert t.shape == (self.n_mbs,)
        _check(_lib.evxa_sao_filter(self.h, slot, t.ctypes.data))

    def deblock(self, slot, records=None):
        """Deblock picture slot `slot` in place (the block types and q_index
        of `records`, n_mbs block records; None: the last analysed frame's)."""
        ptr = None
        if records is not None:
            records = np.ascontiguousarray(records, RECORD_DTYPE)
            assert records.shape == (self.n_mbs,)
            ptr = records.ctypes.data
        _check(_lib.evxa_deblock(self.h, slot, ptr))

    def wait(self):
        """Block until this stream's queued jobs are done (other streams'
        work is not waited for)."""
        _check(_lib.evxa_stream_wait(self.h))

    def query(self) -> bool:
        done = C.c_int32()
        _check(_lib.evxa_stream_query(self.h, C.byref(done)))
        return bool(done.value)

    def reset_temporal(self):
        _check(_lib.evxa_stream_reset_temporal(self.h))

    def _note(self, desc):
        pass

    def analyze(self, desc: FrameDesc):
        _check(_lib.evxa_analyze_frame(self.h, C.byref(desc)))

    def analyze_noalloc(self, desc: FrameDesc):
        """evxa_analyze_frame_noalloc: the wavefront analysis with its launch
        scratch in the library's per-GPU array (no allocation)."""
        _check(_lib.evxa_analyze_frame_noalloc(self.h, C.byref(desc)))

    def encode(self, desc: EncodeDesc):
        """evxa_encode_frame: the in-loop encode of one frame into
        desc.dst_slot; results() then returns its descriptors, coefficients
        and post-trellis levels."""
        _check(_lib.evxa_encode_frame(self.h, C.byref(desc)))

    def encode_frame(self, ftype, quality, src_slot, dst_slot, ref_slots=(), **kw):
        self.encode(encode_desc(ftype, quality, src_slot, dst_slot, ref_slots, **kw))

    def analyze_p(self, quality, src_slot, ref_slots, **kw):
        self.analyze(frame_desc(FRAME_P, quality, src_slot, ref_slots, **kw))

    def analyze_b(self, quality, src_slot, fwd_slot, bwd_slot, **kw):
        self.analyze(frame_desc(FRAME_B, quality, src_slot, [fwd_slot, bwd_slot], **kw))

    def results(self, records=True, coeffs=True, levels=True):
        rec = np.zeros(self.n_mbs, RECORD_DTYPE) if records else None
        co = np.zeros((self.n_mbs, MB_SAMPLES), np.int16) if coeffs else None
        lv = np.zeros((self.n_mbs, MB_SAMPLES), np.int16) if levels else None
        _check(_lib.evxa_read_results(
            self.h, None if rec is None else rec.ctypes.data_as(C.c_void_p),
            None if co is None else _p16(co), None if lv is None else _p16(lv)))
        return rec, co, lv

    def interpolate_subpixel_6tap(self, slot, xs, ys, dir_x, dir_y, quarter):
        xs, ys, dx, dy, q = (np.ascontiguousarray(np.atleast_1d(a), np.int32)
                             for a in (xs, ys, dir_x, dir_y, quarter))
        n = len(xs)
        out = np.empty((n, MB_SAMPLES), np.int16)
        _check(_lib.evxa_interp_blocks(self.h, slot, n, _p32(xs), _p32(ys), _p32(dx), _p32(dy),
                                       _p32(q), _p16(out)))
        return out

    def perform_inter_search(self, src_slot, ref_slot, quality, xs, ys, pred_x, pred_y):
        xs, ys = (np.ascontiguousarray(np.atleast_1d(a), np.int32) for a in (xs, ys))
        px, py = (np.ascontiguousarray(np.atleast_1d(a), np.int16) for a in (pred_x, pred_y))
        n = len(xs)
        out = np.zeros(n, SEARCH_DTYPE)
        _check(_lib.evxa_search_blocks(self.h, src_slot, ref_slot, quality, n, _p32(xs),
                                       _p32(ys), _p16(px), _p16(py),
                                       out.ctypes.data_as(C.c_void_p)))
        return out

    def intra_search(self, src_slot, pic_slot, quality, xs, ys):
        """calculate_intra_prediction of the MBs at (xs, ys) of src_slot over
        pic_slot as the picture being reconstructed."""
        xs, ys = (np.ascontiguousarray(np.atleast_1d(a), np.int32) for a in (xs, ys))
        out = np.zeros(len(xs), SEARCH_DTYPE)
        _check(_lib.evxa_intra_search_blocks(self.h, src_slot, pic_slot, quality, len(xs),
                                             _p32(xs), _p32(ys),
                                             out.ctypes.data_as(C.c_void_p)))
        return out


def clock() -> float:
    """The dispatcher's monotonic clock (seconds)."""
    return load_library().evxa_clock()


class Server:
    """Continuous batching: a native dispatcher keeps up to `max_inflight`
    fused launches on the device and fuses every frame that is ready (its
    stream idle) into the next free one, up to `max_frames` per launch."""

    def __init__(self, ctx: Context, max_inflight: int = 2, max_frames: int = 64):
        self.ctx = ctx
        h = C.c_void_p()
        _check(_lib.evxa_server_create(ctx.h, max_inflight, max_frames, C.byref(h)))
        self.h = h
        self._buf = np.zeros(4096, COMPLETION_DTYPE)

    def submit(self, stream: "Stream", desc: FrameDesc, tag: int, t_arrival: float = 0.0):
        _check(_lib.evxa_server_submit(self.h, stream.h, C.byref(desc), tag, t_arrival))

    def poll(self):
        """Completions available now (structured array of COMPLETION_DTYPE)."""
        out = []
        while True:
            n = C.c_int32()
            _check(_lib.evxa_server_poll(self.h, self._buf.ctypes.data_as(C.c_void_p),
                                         len(self._buf), C.byref(n)))
            out.append(self._buf[:n.value].copy())
            if n.value < len(self._buf):
                return np.concatenate(out)

    def pending(self):
        q, f = C.c_int32(), C.c_int32()
        _check(_lib.evxa_server_pending(self.h, C.byref(q), C.byref(f)))
        return q.value, f.value

    def close(self):
        if self.h:
            _check(_lib.evxa_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
