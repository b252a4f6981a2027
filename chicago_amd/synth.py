"""Seeded synthetic 8-bit 4:2:0 content for the analysis workloads.

Frames are int16 planes holding 8-bit samples, aligned up to 16 with zero
rows the way the reference pads its images (evx3enc.cpp:93-94,
image.cpp:89).  Generated once on the host; the CPU and GPU paths consume the
same buffers, so libm differences never reach a parity check.

Recipes
  texture:  Y = clamp(round(128 + sum_k (40/6) sin(2 pi (fx_k x + fy_k y) + phi_k)
                 + U[-8, 8]), 0, 255), k = 1..6, f uniform in +-1/32 cycles/px.
            Chroma planes are independent draws of the same recipe at half size.
  moved:    every 16x16 block (with its 8x8 chroma) is fetched from the previous
            picture at an integer displacement (clamped inside the picture) and
            an optional half/quarter-pel state that is sampled with the 6-tap
            rule of interp.cpp:52-164, then Gaussian noise (sigma 2) is added
            and the result is clamped to [0, 255].
"""
from __future__ import annotations

import numpy as np

DIRS = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]


def align16(n: int) -> int:
    return (n + 15) & ~15


def seed_of(*parts) -> int:
    """Stable 64-bit seed from small integers (config, stream, purpose...)."""
    h = 0x9E3779B97F4A7C15
    for p in parts:
        h ^= (int(p) + 0x9E3779B97F4A7C15 + ((h << 6) & 0xFFFFFFFFFFFFFFFF) + (h >> 2))
        h &= 0xFFFFFFFFFFFFFFFF
    return h


def _texture_plane(rng, w, h):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    acc = np.full((h, w), 128.0)
    for _ in range(6):
        fx, fy = rng.uniform(-1 / 32, 1 / 32, size=2)
        phi = rng.uniform(0, 2 * np.pi)
        acc += (40.0 / 6.0) * np.sin(2 * np.pi * (fx * xx + fy * yy) + phi)
    acc += rng.integers(-8, 9, size=(h, w))
    return np.clip(np.rint(acc), 0, 255).astype(np.int16)


def texture_frame(rng, width, height):
    """(y, u, v) at the aligned size; rows past `height` are zero."""
    aw, ah = align16(width), align16(height)
    y = np.zeros((ah, aw), np.int16)
    u = np.zeros((ah // 2, aw // 2), np.int16)
    v = np.zeros((ah // 2, aw // 2), np.int16)
    y[:height, :width] = _texture_plane(rng, width, height)
    u[:height // 2, :width // 2] = _texture_plane(rng, width // 2, height // 2)
    v[:height // 2, :width // 2] = _texture_plane(rng, width // 2, height // 2)
    return y, u, v


def _clamped(p, xs, ys):
    h, w = p.shape
    return p[np.clip(ys, 0, h - 1), np.clip(xs, 0, w - 1)].astype(np.int32)


def _tap6(p, xs, ys, axis_x):
    """6-tap half sample at (xs + 0.5, ys) if axis_x else (xs, ys + 0.5)."""
    taps = (1, -5, 20, 20, -5, 1)
    acc = 0
    for k, t in enumerate(taps):
        o = k - 2
        acc = acc + t * (_clamped(p, xs + o, ys) if axis_x else _clamped(p, xs, ys + o))
    return ((acc + 16) >> 5).astype(np.int16)


def subpel_plane(p, dx, dy, quarter):
    """Whole-plane sub-pel samples: value at every integer base position of
    the 6-tap rule (interp.cpp:52-164).  Depends only on absolute position."""
    h, w = p.shape
    ys, xs = np.mgrid[0:h, 0:w]
    xo = 0 if dx > 0 else -1
    yo = 0 if dy > 0 else -1
    if dx != 0 and dy == 0:
        half = _tap6(p, xs + xo, ys, True)
    elif dx == 0 and dy != 0:
        half = _tap6(p, xs, ys + yo, False)
    else:
        taps = (1, -5, 20, 20, -5, 1)
        acc = 0
        for k, t in enumerate(taps):
            acc = acc + t * _tap6(p, xs + xo, ys + yo + k - 2, True).astype(np.int32)
        half = ((acc + 16) >> 5).astype(np.int16)
    if quarter:
        return ((p.astype(np.int32) + half.astype(np.int32) + 1) >> 1).astype(np.int16)
    return half


_TAPS = (1, -5, 20, 20, -5, 1)


def _patches(p, x0, y0, size):
    """(n, size, size) int32 patches of plane p with top-left corners
    (x0, y0), coordinates clamped into the plane like interp.cpp:8-14."""
    h, w = p.shape
    r = np.arange(size)
    rows = np.clip(y0[:, None] + r[None, :], 0, h - 1)
    cols = np.clip(x0[:, None] + r[None, :], 0, w - 1)
    return p[rows[:, :, None], cols[:, None, :]].astype(np.int32)


def _subpel_blocks(patch, bs, dx, dy, quarter):
    """subpel_plane's samples for bs x bs blocks whose patches start 3 samples
    above and left of the block (patch size bs + 6)."""
    xo = 0 if dx > 0 else -1
    yo = 0 if dy > 0 else -1

    def hpass(r0, nrows):
        # Horizontal halves between columns c + xo and c + xo + 1 of patch rows
        # r0 .. r0 + nrows - 1 for block columns c = 0 .. bs - 1.
        acc = 0
        for k, t in enumerate(_TAPS):
            c0 = 3 + xo + k - 2
            acc = acc + t * patch[:, r0:r0 + nrows, c0:c0 + bs]
        return ((acc + 16) >> 5).astype(np.int16)

    if dx != 0 and dy == 0:
        half = hpass(3, bs)
    elif dx == 0 and dy != 0:
        acc = 0
        for k, t in enumerate(_TAPS):
            r0 = 3 + yo + k - 2
            acc = acc + t * patch[:, r0:r0 + bs, 3:3 + bs]
        half = ((acc + 16) >> 5).astype(np.int16)
    else:
        rows = hpass(3 + yo - 2, bs + 5).astype(np.int32)
        acc = 0
        for k, t in enumerate(_TAPS):
            acc = acc + t * rows[:, k:k + bs, :]
        half = ((acc + 16) >> 5).astype(np.int16)
    if quarter:
        full = patch[:, 3:3 + bs, 3:3 + bs]
        return ((full + half.astype(np.int32) + 1) >> 1).astype(np.int16)
    return half


def move_blocks(rng, prev, mvx, mvy, sp_state, noise_sigma=2.0, visible=None):
    """Build a picture from `prev` block by block.

    mvx/mvy: int arrays (hb, wb) of integer displacements; sp_state: int array
    (hb, wb) with -1 for none, else dir_index*2 + quarter (dir_index into DIRS).
    Sub-pel samples follow subpel_plane (interp.cpp:52-164), computed only on
    each block's clamped neighbourhood.
    """
    ah, aw = prev[0].shape
    hb, wb = ah // 16, aw // 16
    bys, bxs = np.mgrid[0:hb, 0:wb]
    sx = np.clip(bxs * 16 + np.asarray(mvx), 0, aw - 16).ravel()
    sy = np.clip(bys * 16 + np.asarray(mvy), 0, ah - 16).ravel()
    sp_state = np.broadcast_to(np.asarray(sp_state), (hb, wb)).ravel()
    out = []
    for ci, src in enumerate(prev):
        bs, sh = (16, 0) if ci == 0 else (8, 1)
        blocks = np.empty((hb * wb, bs, bs), np.int16)
        for st in np.unique(sp_state):
            idx = np.nonzero(sp_state == st)[0]
            patch = _patches(src, (sx[idx] >> sh) - 3, (sy[idx] >> sh) - 3, bs + 6)
            if st < 0:
                blocks[idx] = patch[:, 3:3 + bs, 3:3 + bs]
            else:
                blocks[idx] = _subpel_blocks(patch, bs, DIRS[st // 2][0], DIRS[st // 2][1], st & 1)
        out.append(blocks.reshape(hb, wb, bs, bs).transpose(0, 2, 1, 3).reshape(hb * bs, wb * bs))
    for i in range(3):
        n = np.rint(rng.normal(0.0, noise_sigma, size=out[i].shape)) if noise_sigma > 0 else 0
        out[i] = np.clip(out[i].astype(np.float64) + n, 0, 255).astype(np.int16)
    if visible is not None:
        w, h = visible
        out[0][h:, :] = 0
        out[1][h // 2:, :] = 0
        out[2][h // 2:, :] = 0
    return tuple(out)


def pair_c0(seed, width=1280, height=720, mv_range=16):
    """Reference + current picture: random integer MV in [-16, 16]^2 and a
    random sub-pel state per block (none, or one of 8 dirs x half/quarter)."""
    rng = np.random.default_rng(seed)
    ref = texture_frame(rng, width, height)
    ah, aw = ref[0].shape
    hb, wb = ah // 16, aw // 16
    mvx = rng.integers(-mv_range, mv_range + 1, size=(hb, wb))
    mvy = rng.integers(-mv_range, mv_range + 1, size=(hb, wb))
    sp = rng.integers(-1, 16, size=(hb, wb))
    cur = move_blocks(rng, ref, mvx, mvy, sp, visible=(width, height))
    return ref, cur


def sequence_c1(seed, n_frames, width=1920, height=1080, local_range=24, local_frac=0.3,
                pan_quarter=True):
    """Frame 0 from the texture recipe; frame t is frame t-1 moved by a global
    pan of 1.25 px/frame (integer 1 plus the quarter-pel (-1, 0) state) with
    uniform +-`local_range` px local motion on `local_frac` of the blocks."""
    rng = np.random.default_rng(seed)
    frames = [texture_frame(rng, width, height)]
    ah, aw = frames[0][0].shape
    hb, wb = ah // 16, aw // 16
    st = DIRS.index((-1, 0)) * 2 + 1 if pan_quarter else -1
    for _ in range(1, n_frames):
        mvx = np.full((hb, wb), -1)
        mvy = np.zeros((hb, wb), np.int64)
        local = rng.random((hb, wb)) < local_frac
        mvx = mvx + np.where(local, rng.integers(-local_range, local_range + 1, (hb, wb)), 0)
        mvy = mvy + np.where(local, rng.integers(-local_range, local_range + 1, (hb, wb)), 0)
        sp = np.full((hb, wb), st)
        frames.append(move_blocks(rng, frames[-1], mvx, mvy, sp, visible=(width, height)))
    return frames


def _subpel_state(fx, fy):
    """Nearest representable sub-pel state for a fractional offset in
    [-0.5, 0.5]^2: none, or a direction with a half (0.5) or quarter (0.25)
    step along it (the states of move_blocks)."""
    ax, ay = abs(fx), abs(fy)
    if max(ax, ay) < 0.125:
        return -1
    dx = 0 if ax < 0.125 else (1 if fx > 0 else -1)
    dy = 0 if ay < 0.125 else (1 if fy > 0 else -1)
    quarter = 1 if max(ax, ay) < 0.375 else 0
    return DIRS.index((dx, dy)) * 2 + quarter


def sequence_c2(seed, n_frames=33, width=3840, height=2160, zoom=0.005, local_range=32,
                local_frac=0.3):
    """Frame 0 from the texture recipe; frame t is frame t-1 zoomed in by
    `zoom` about the picture centre (each block fetched from where the zoom
    maps its centre, rounded to the nearest integer plus half/quarter-pel
    state) with uniform +-`local_range` px local motion on `local_frac` of
    the blocks."""
    rng = np.random.default_rng(seed)
    frames = [texture_frame(rng, width, height)]
    ah, aw = frames[0][0].shape
    hb, wb = ah // 16, aw // 16
    bys, bxs = np.mgrid[0:hb, 0:wb]
    cx, cy = (width - 1) / 2.0, (height - 1) / 2.0
    # Source position of a block centre p under a zoom-in by s: c + (p - c) / s.
    dxf = ((bxs * 16 + 7.5) - cx) * (1.0 / (1.0 + zoom) - 1.0)
    dyf = ((bys * 16 + 7.5) - cy) * (1.0 / (1.0 + zoom) - 1.0)
    ix, iy = np.rint(dxf).astype(np.int64), np.rint(dyf).astype(np.int64)
    sp_zoom = np.vectorize(_subpel_state)(dxf - ix, dyf - iy)
    for _ in range(1, n_frames):
        local = rng.random((hb, wb)) < local_frac
        mvx = ix + np.where(local, rng.integers(-local_range, local_range + 1, (hb, wb)), 0)
        mvy = iy + np.where(local, rng.integers(-local_range, local_range + 1, (hb, wb)), 0)
        frames.append(move_blocks(rng, frames[-1], mvx, mvy, sp_zoom, visible=(width, height)))
    return frames


def blocky_picture(seed, width, height, amp=8, noise=2, base=120):
    """A smooth picture with a random offset per 8x8 block and light noise:
    the blocking a coarse quantiser leaves, for the deblocking filter.
    Returns (y, u, v) int16 planes (chroma width/2 x height/2)."""
    rng = np.random.default_rng(seed)

    def plane(w, h):
        yy, xx = np.mgrid[0:h, 0:w]
        smooth = base + 40 * np.sin(xx / 37.0) + 30 * np.cos(yy / 23.0)
        off = rng.integers(-amp, amp + 1, (h // 8 + 1, w // 8 + 1))
        steps = np.repeat(np.repeat(off, 8, 0), 8, 1)[:h, :w]
        return (smooth + steps + rng.integers(-noise, noise + 1, (h, w))).astype(np.int16)

    return plane(width, height), plane(width // 2, height // 2), plane(width // 2, height // 2)


def random_block_records(seed, width, height, q_lo=1, q_hi=63, copy_frac=0.25):
    """Block records for a width x height frame with random block types
    (delta or copy, with or without motion) and q_index in [q_lo, q_hi]."""
    from .evxa import RECORD_DTYPE
    rng = np.random.default_rng(seed)
    n = (width // 16) * (height // 16)
    rec = np.zeros(n, RECORD_DTYPE)
    copy = rng.random(n) < copy_frac
    rec["block_type"] = (copy * 4) | (rng.integers(0, 2, n) * 2)
    rec["q_index"] = rng.integers(q_lo, q_hi + 1, n)
    return rec
