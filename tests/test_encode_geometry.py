"""Geometry of the in-loop encode kernel, checked on the CPU: the window of
the picture being reconstructed that each MB stages holds every read its
intra search, intra predictions and intra block copies can make, and lies
inside the picture's allocation (interior plus a 24-sample apron); the MB
order puts every MB after the MBs it waits on and reads.  The constants
restate encode_mb's and encode_order_kernel's (chicago_amd/csrc)."""
import pytest

APRON = 24
LAG = 3                      # MB (bx, by) after (bx + 3, by - 1)
LUMA = (84, 84)              # window columns, rows
CHROMA = (46, 45)


def pitch(w):
    return ((w + 2 * APRON + 63) // 64) * 64


def window(W, H, x, y):
    x0 = min(max(x - 34, -APRON), W + APRON - LUMA[0])
    y0 = min(max(y - 50, -APRON), H + APRON - LUMA[1])
    cx0 = min(max((x >> 1) - 20, -APRON), (W >> 1) + APRON - CHROMA[0])
    cy0 = min(max((y >> 1) - 27, -APRON), (H >> 1) + APRON - CHROMA[1])
    return x0, y0, cx0, cy0


def check_frame(W, H):
    for y in range(0, H, 16):
        for x in range(0, W, 16):
            x0, y0, cx0, cy0 = window(W, H, x, y)
            # dword-aligned rows, inside the allocation
            assert x0 % 2 == 0 and cx0 % 2 == 0
            assert x0 >= -APRON and x0 + LUMA[0] <= pitch(W) - APRON
            assert y0 >= -APRON and y0 + LUMA[1] <= H + APRON
            assert cx0 >= -APRON and cx0 + CHROMA[0] <= pitch(W // 2) - APRON
            assert cy0 >= -APRON and cy0 + CHROMA[1] <= H // 2 + APRON
            # candidates: within 15 of the first 3x3 scan (x - 16 .. x + 16,
            # y - 32 .. y), clamped into the picture; 6-tap reads 3 before and
            # 18 after (luma), 3 before and 10 after (chroma); block loads
            # read up to 2 samples past a row's last 4-sample group.
            cxs = range(max(0, x - 31), min(W - 16, x + 31) + 1)
            cys = range(max(0, y - 47), min(H - 16, y + 15) + 1)
            assert min(cxs) - 3 >= x0 and max(max(cxs) + 18, max(cxs) + 17) < x0 + LUMA[0]
            assert min(cys) - 3 >= y0 and max(cys) + 18 < y0 + LUMA[1]
            ccx = [c >> 1 for c in cxs]
            ccy = [c >> 1 for c in cys]
            assert min(ccx) - 3 >= cx0 and max(ccx) + 10 < cx0 + CHROMA[0]
            assert min(ccy) - 3 >= cy0 and max(ccy) + 10 < cy0 + CHROMA[1]
            # neighbours of the directional intra modes
            assert y - 1 >= y0 and x - 1 >= x0 and (y >> 1) - 1 >= cy0 and (x >> 1) - 1 >= cx0


def test_window_holds_every_read():
    for W in range(48, 420, 16):
        for H in range(48, 260, 16):
            check_frame(W, H)
    for W, H in ((1280, 720), (1920, 1088), (2560, 1440), (3840, 2160)):
        check_frame(W, H)


def test_unwritten_rows_below_stay_unwritten():
    """A MB reads rows below its own only in columns x - 34 .. x + 2 (a
    candidate left of x - 16); under the lag the row below has at most
    reached MB bx - LAG - 1, which ends before x - 34."""
    for bx in range(LAG + 1, 40):
        x = 16 * bx
        done_end = 16 * (bx - LAG - 1) + 15
        assert done_end < x - 34
    # and the row above is done far enough: reads reach column x + 49
    for bx in range(0, 40):
        assert 16 * (bx + LAG) + 15 >= 16 * bx + 49


def order(wb, hb):
    slope = LAG + 1
    out = [None] * (wb * hb)
    for i in range(wb * hb):
        bx, by = i % wb, i // wb
        v = bx + slope * by
        before = 0
        for yy in range(hb):
            c = v - slope * yy
            before += 0 if c <= 0 else (c if c < wb else wb)
            if yy < by and 0 <= c < wb:
                before += 1
        assert out[before] is None
        out[before] = i
    return out


@pytest.mark.parametrize("wb,hb", [(1, 1), (3, 7), (8, 6), (20, 11), (120, 68)])
def test_order_respects_dependencies(wb, hb):
    out = order(wb, hb)
    assert sorted(out) == list(range(wb * hb))
    pos = {m: k for k, m in enumerate(out)}
    for i in range(wb * hb):
        bx, by = i % wb, i // wb
        if bx > 0:
            assert pos[i - 1] < pos[i]
        if by > 0:
            for dx in range(-1, LAG + 1):
                if 0 <= bx + dx < wb:
                    assert pos[(by - 1) * wb + bx + dx] < pos[i]
