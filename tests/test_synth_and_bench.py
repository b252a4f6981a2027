"""Synthetic content generator and the bench.py contract.  CPU only."""
import json
import os
import subprocess
import sys

import numpy as np

from chicago_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_alignment_and_zero_rows():
    ref, cur = synth.pair_c0(3, 640, 360)
    for pic in (ref, cur):
        y, u, v = pic
        assert y.shape == (368, 640) and u.shape == (184, 320) and v.shape == (184, 320)
        assert not y[360:].any() and not u[180:].any() and not v[180:].any()
        assert y[:360].min() >= 0 and y[:360].max() <= 255


def test_deterministic():
    a = synth.sequence_c1(5, 3, 96, 64)
    b = synth.sequence_c1(5, 3, 96, 64)
    for fa, fb in zip(a, b):
        for pa, pb in zip(fa, fb):
            assert np.array_equal(pa, pb)


def test_subpel_plane_constant():
    p = np.full((20, 30), 77, np.int16)
    for dx, dy in synth.DIRS:
        for q in (0, 1):
            assert (synth.subpel_plane(p, dx, dy, q) == 77).all()


def test_pan_moves_content():
    """Frame t is frame t-1 shifted right by 1.25 px (noise aside): the
    integer part of the motion search finds the pan."""
    frames = synth.sequence_c1(11, 2, 128, 64, local_frac=0.0)
    a = frames[0][0].astype(float)
    b = frames[1][0].astype(float)
    err = [np.abs(b[8:56, 16:112] - a[8:56, 16 - s:112 - s]).mean() for s in range(0, 3)]
    assert int(np.argmin(err)) == 1


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
