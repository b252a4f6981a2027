"""The intra-frame block search (calculate_intra_prediction, motion.cpp:485-565)
on the GPU through evxa_intra_search_blocks, against the oracle: every MB
(edges included) of natural, blocky, tiled (block copies) and 16-bit
content, every record field including mad, cost and the candidate counters."""
import numpy as np
import pytest

from chicago_amd import evxa
from oracle.oracle import Picture

from test_oracle_vs_reference import INTRA_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", range(len(INTRA_CASES)), ids=[c[0] for c in INTRA_CASES])
@pytest.mark.parametrize("quality", [0, 8, 13, 63])
def test_intra_search_matches_oracle(ctx, oracle, case, quality):
    _, src, pic = INTRA_CASES[case]
    h, w = src[0].shape
    s = ctx.open_stream(w, h, 2)
    s.upload(0, *src)
    s.upload(1, *pic)
    S, P = Picture(*src), Picture(*pic)
    ys, xs = np.mgrid[0:h:16, 0:w:16]
    xs, ys = xs.ravel(), ys.ravel()
    got = s.intra_search(0, 1, quality, xs, ys)
    for k, (x, y) in enumerate(zip(xs, ys)):
        r = oracle.intra_search(S.block(int(x), int(y)), P, int(x), int(y), quality)
        want = tuple(getattr(r, f) for f in evxa.SEARCH_DTYPE.names)
        assert tuple(int(got[k][f]) for f in evxa.SEARCH_DTYPE.names) == want, (int(x), int(y))
    s.close()


def test_intra_search_rejects_bad_arguments(ctx):
    s = ctx.open_stream(64, 48, 2)
    with pytest.raises(evxa.EvxaError):
        s.intra_search(0, 1, 64, [0], [0])
    with pytest.raises(evxa.EvxaError):
        s.intra_search(0, 1, 8, [56], [0])
    with pytest.raises(evxa.EvxaError):
        s.intra_search(0, 2, 8, [0], [0])
    s.close()
