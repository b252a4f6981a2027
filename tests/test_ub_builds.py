"""UBSan and -fwrapv builds of the oracle and of the reference's sources
(oracle/ub_check.py): on the golden fixtures' inputs and random full-range
blocks the oracle has no undefined operation and no result that depends on
how signed overflow is compiled; the reference overflows in exactly two
places, both in compute_block_variance2 (analysis.h:269, the sum of squares,
and analysis.h:274, the square of the sum), and only variance2 changes
under -fwrapv -- the golden vectors hold the -O3 build's values, which the
oracle and the kernels follow.  Skipped where the reference tree is absent.
CPU only."""
import os
import subprocess

import pytest

from oracle import ub_check

REF_SRC = "/root/reference/src"
pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference tree absent")


@pytest.fixture(scope="module")
def report():
    subprocess.run(["make", "-s", "-C", ub_check.HERE, "oracle", "ref", "checks", "-j8"],
                   check=True, timeout=600)
    assert ub_check.available()
    res, text = ub_check.report()
    print(text)
    return res


def test_oracle_has_no_undefined_behaviour(report):
    assert report["ubsan", "oracle"][1] == {}


def test_reference_overflows_only_in_variance2(report):
    assert sorted(report["ubsan", "ref"][1]) == [("video/analysis.h", 269),
                                                 ("video/analysis.h", 274)]


def test_wrap_semantics_change_only_variance2(report):
    assert ub_check.differing(report["O3", "oracle"][0], report["fwrapv", "oracle"][0]) == []
    assert ub_check.differing(report["O3", "ref"][0], report["fwrapv", "ref"][0]) == ["variance2"]
    assert ub_check.differing(report["O3", "ref"][0], report["ubsan", "ref"][0]) == ["variance2"]


def test_oracle_equals_reference_build_on_every_workload(report):
    a, b = report["O3", "oracle"][0], report["O3", "ref"][0]
    assert ub_check.differing(a, b) == []
    assert len(set(a) & set(b)) >= 40
