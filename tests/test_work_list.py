"""The device-built work list of the allocation-free path lists a frame's
MBs in the wavefront's anti-diagonal order (v = bx + 2 by, a diagonal's MBs
by row, a MB's references adjacent).  Its closed-form position of each MB
(the MBs before it) is checked here against a brute-force enumeration of
that order."""
import pytest


def closed_form_position(bx, by, wb, hb):
    # wave_list_kernel's count: MBs (x, y) with x + 2y < v, plus those of
    # diagonal v on earlier rows.
    v = bx + 2 * by
    before = 0
    for y in range(hb):
        c = v - 2 * y
        before += 0 if c <= 0 else min(c, wb)
        if y < by and 0 <= c < wb:
            before += 1
    return before


def enumerated_order(wb, hb):
    order = []
    for v in range(wb + 2 * (hb - 1)):
        for by in range(hb):
            bx = v - 2 * by
            if bx < 0:
                break
            if bx < wb:
                order.append((bx, by))
    return order


@pytest.mark.parametrize("wb,hb", [(1, 1), (1, 5), (5, 1), (2, 3), (8, 6), (40, 23), (120, 68)])
def test_positions_match_enumeration(wb, hb):
    order = enumerated_order(wb, hb)
    assert len(order) == wb * hb
    for pos, (bx, by) in enumerate(order):
        assert closed_form_position(bx, by, wb, hb) == pos, (bx, by)
