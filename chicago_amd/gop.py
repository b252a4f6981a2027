"""Frame roles of the hierarchical-B workload: GOP 8 with I/P anchors and a
3-level B pyramid, in decode order.

The reference's own GOP is one P anchor plus a 2-level B hierarchy of 3
frames (evx3enc.cpp:279-383): the anchor references the previous anchor, the
middle (reference) B references (previous anchor, new anchor) at quality + 1,
and the other (non-reference) Bs reference the nearest decoded pictures on
either side at quality + 2, capped at 31 (evx3enc.cpp:318, 337).  A GOP of 8
with a 3-level pyramid cannot be expressed with EVX_B_FRAME_COUNT <= 3
(config.h:121), so this schedule extends the same rule: every B references
the nearest already-decoded pictures before and after it; Bs that later
frames reference (levels 1 and 2) use quality + 1, the rest quality + 2.
P anchors use one reference (the previous anchor) and the temporal MV table
of the previous P anchor (encode.cpp:990-997 copies the table after every
non-B frame; after the I frame it holds no motion blocks, which equals no
temporal candidate).
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
