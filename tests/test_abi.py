"""The C-ABI library: it loads without a GPU, exports every symbol that
include/evxa.h declares, and refuses to run without a device (there is no
CPU fallback).  CPU only: no compute call is made."""
import ctypes
import os
import re

import pytest

from chicago_amd import evxa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "evxa.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(evxa_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(evxa.EXPORTS)


@pytest.mark.skipif(not os.path.exists(evxa.LIB_PATH), reason="libevxa.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(evxa.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(evxa.LIB_PATH), reason="libevxa.so not built")
def test_host_only_entry_points():
    lib = evxa.load_library()
    assert lib.evxa_version() >= 100
    # Stream footprint used for HBM admission grows with size and slots.
    a = lib.evxa_stream_bytes(1920, 1088, 5)
    assert a > 5 * 1920 * 1088 * 3  # five int16 4:2:0 pictures at least
    assert lib.evxa_stream_bytes(1920, 1088, 6) > a
    assert lib.evxa_stream_bytes(0, 1088, 5) == 0


def test_records_layout():
    assert evxa.RECORD_DTYPE.itemsize == 64
    assert evxa.RECORD_DTYPE.fields["sad"][1] == 24
    assert evxa.RECORD_DTYPE.fields["flags"][1] == 56
    from oracle.oracle import RECORD_DTYPE
    assert RECORD_DTYPE == evxa.RECORD_DTYPE


@pytest.mark.skipif(not os.path.exists(evxa.LIB_PATH), reason="libevxa.so not built")
def test_no_device_fails_loudly():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(evxa.EvxaError):
        evxa.Context(0)
