"""Frame roles in decode order: the reference encoder's own GOP, and the
GOP-8 hierarchical-B workload.

reference_schedule restates the reference encoder's frame ordering
(evx3enc.cpp:104-239, 279-415) with its picture-buffer slots
(common.cpp:263-281): one P anchor plus a 2-level hierarchy of up to 3 Bs.
The anchor references the previous anchor; the middle (reference) B
references (previous anchor, new anchor) at quality + 1, and the other
(non-reference) Bs reference the nearest decoded pictures on either side at
quality + 2, capped at 31 (evx3enc.cpp:318, 337).

gop8_schedule extends the same rule to a GOP of 8 with a 3-level pyramid,
which EVX_B_FRAME_COUNT <= 3 (config.h:121) cannot express: every B
references the nearest already-decoded pictures before and after it; Bs that
later frames reference (levels 1 and 2) use quality + 1, the rest quality
+ 2.  On whole GOPs of 4 it gives the reference's roles.

P anchors use one reference (the previous anchor) and, as temporal MV table,
the block table of the previous non-B frame (encode.cpp:990-997 copies the
table after every non-B frame).  After an I frame that table comes from the
intra analysis, which this path does not include: the schedules' users start
the first P after an I with an empty table (gop8_schedule marks it
temporal_mvp=False), as the reference does at the start of a stream, where
the table is zero-filled (common.cpp:175-183).
"""
from __future__ import annotations

from dataclasses import dataclass, field


@dataclass
class FrameRole:
    index: int                 # display index
    kind: str                  # "I", "P" or "B"
    refs: list = field(default_factory=list)   # display indices: P [prev anchor], B [fwd, bwd]
    quality: int = 0
    level: int = 0             # B pyramid level (1..3), 0 for anchors
    temporal_mvp: bool = False  # P: use the previous P anchor's table
    # Decoded-picture-buffer placement (reference_schedule only): the slot the
    # reconstruction goes to, the slots the references are read from (they
    # are the records' prediction_target values), whether later frames
    # reference this one, and the display index of the frame whose block
    # table is the temporal MV table (P frames; -1 for none).
    slot: int = -1
    ref_slots: list = field(default_factory=list)
    is_reference: bool = True
    temporal_from: int = -1


# The reference encoder's configuration (config.h:41, 121; quantize.h:42).
B_FRAME_COUNT = 3
PERIODIC_INTRA_RATE = 3600
MAX_QUALITY = 63


def internal_quality(user_quality: int) -> int:
    """set_quality (evx3enc.cpp:64-82): the user's 1..31 mapped to the QP the
    frames carry, unchanged up to 16, then q + (q - 16)^2 / 10, capped at 63."""
    q = min(max(int(user_quality), 1), 31)
    qp = q if q <= 16 else q + ((q - 16) * (q - 16)) // 10
    return min(qp, MAX_QUALITY)


def assign_ip_dpb_slot(ip_index: int) -> int:
    """common.cpp:263-266: I/P anchors alternate between slots 0 and 1 (the
    index is a uint32 there, so -1 wraps to an odd number: slot 1)."""
    return (ip_index & 0xFFFFFFFF) % 2


def assign_bref_dpb_slot() -> int:
    """common.cpp:268-271: the reference B lives in slot 2."""
    return 2


def assign_b_scratch_slot_for_count(b_frame_count: int) -> int:
    """common.cpp:278-281: non-reference Bs go to slot 3 when the GOP has a
    reference B (3 Bs), else to slot 2."""
    return 3 if b_frame_count >= 3 else 2


def assign_b_scratch_slot() -> int:
    """common.cpp:273-276, at the configured B count."""
    return assign_b_scratch_slot_for_count(B_FRAME_COUNT)


def reference_schedule(n_frames: int, quality: int, b_count: int = B_FRAME_COUNT,
                       intra_rate: int = PERIODIC_INTRA_RATE):
    """Decode-order roles the reference encoder gives display frames
    0 .. n_frames-1 fed one at a time and then flushed: encode()
    (evx3enc.cpp:104-239), encode_gop (:279-383) and flush (:385-415).

    `quality` is the frame QP (internal_quality of the user's setting).  The
    first frame, and every `intra_rate`-th, is an I frame; frames are then
    buffered until b_count + 1 are pending, and the last pending frame is
    coded first as a P anchor referencing the previous anchor's slot.  With
    Bs pending and b_count >= 3 configured, the B at b_count / 2 (of those
    pending) is a reference B at min(quality + 1, 31) between the two
    anchors, and the other Bs are non-reference at min(quality + 2, 31),
    each between the reference B and the nearer anchor; with fewer Bs
    configured, every B sits between the two anchors at quality + 2.  P
    frames take the block table of the previous non-B frame as their
    temporal MV table (encode.cpp:990-997)."""
    roles = []
    pending = []
    state = {"ip": 0, "last_nonb": -1}

    def emit(role):
        roles.append(role)
        if role.kind != "B":
            state["last_nonb"] = role.index

    def anchor_p(disp):
        ip = state["ip"]
        fwd = assign_ip_dpb_slot(ip - 1)
        prev_anchor = next(r.index for r in reversed(roles) if r.kind != "B")
        emit(FrameRole(disp, "P", [prev_anchor], quality, 0, True, assign_ip_dpb_slot(ip),
                       [fwd], True, state["last_nonb"]))
        return prev_anchor

    def encode_gop():
        ip = state["ip"]
        p_disp = pending[-1]
        prev_anchor = anchor_p(p_disp)
        n_b = len(pending) - 1
        if n_b > 0 and b_count >= 3:
            ref_b = n_b // 2
            bq = min(quality + 1, 31)
            emit(FrameRole(pending[ref_b], "B", [prev_anchor, p_disp], bq, 1, False,
                           assign_bref_dpb_slot(),
                           [assign_ip_dpb_slot(ip - 1), assign_ip_dpb_slot(ip)], True))
            for bi in range(n_b):
                if bi == ref_b:
                    continue
                q2 = min(quality + 2, 31)
                if bi < ref_b:
                    refs = [prev_anchor, pending[ref_b]]
                    slots = [assign_ip_dpb_slot(ip - 1), assign_bref_dpb_slot()]
                else:
                    refs = [pending[ref_b], p_disp]
                    slots = [assign_bref_dpb_slot(), assign_ip_dpb_slot(ip)]
                emit(FrameRole(pending[bi], "B", refs, q2, 2, False,
                               assign_b_scratch_slot_for_count(b_count), slots, False))
        else:
            for bi in range(n_b):
                emit(FrameRole(pending[bi], "B", [prev_anchor, p_disp], min(quality + 2, 31), 1,
                               False, assign_b_scratch_slot_for_count(b_count),
                               [assign_ip_dpb_slot(ip - 1), assign_ip_dpb_slot(ip)], False))
        state["ip"] = ip + 1
        pending.clear()

    force_intra = False
    for disp in range(n_frames):
        if intra_rate and disp > 0 and disp % intra_rate == 0:
            if pending:
                encode_gop()
            force_intra = True
            state["ip"] = 0
        pending.append(disp)
        if not roles or force_intra:
            force_intra = False
            emit(FrameRole(pending[0], "I", [], quality, 0, False,
                           assign_ip_dpb_slot(state["ip"]), [], True))
            state["ip"] += 1
            pending.clear()
        elif len(pending) == b_count + 1:
            encode_gop()
    if len(pending) == 1:
        anchor_p(pending[0])
        state["ip"] += 1
        pending.clear()
    elif pending:
        encode_gop()
    return roles


def gop8_schedule(n_frames: int, quality: int, gop: int = 8):
    """Decode-order roles of frames 0 .. n_frames-1 (I at 0, P every `gop`)."""
    roles = [FrameRole(0, "I", [], quality)]
    prev_anchor = 0
    have_p = False
    while prev_anchor + 1 < n_frames:
        anchor = min(prev_anchor + gop, n_frames - 1)
        roles.append(FrameRole(anchor, "P", [prev_anchor], quality, 0, have_p))
        have_p = True

        def split(lo, hi, level):
            if hi - lo < 2:
                return
            mid = (lo + hi) // 2
            referenced = hi - lo > 2
            q = min(quality + (1 if referenced else 2), 31)
            roles.append(FrameRole(mid, "B", [lo, hi], q, level))
            split(lo, mid, level + 1)
            split(mid, hi, level + 1)

        split(prev_anchor, anchor, 1)
        prev_anchor = anchor
    return roles
