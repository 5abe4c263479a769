"""The C ABI: every entry point include/wavehip.h declares is exported by
libwavehip.so (no device calls -- this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wv_[a-z0-9_]+)\s*\(", src)))


def test_header_lists_match_python_mirror():
    import wavehip
    assert declared("wavehip.h") == sorted(wavehip.EXPORTS)


def test_library_exports_every_declared_symbol():
    import wavehip
    if not os.path.exists(wavehip.LIB_PATH):
        pytest.fail("libwavehip.so not built: run __graft_entry__.build()")
    L = ctypes.CDLL(wavehip.LIB_PATH)
    missing = [s for s in declared("wavehip.h") if not hasattr(L, s)]
    assert not missing


def test_oracle_library_exports():
    import pyoracle
    L = pyoracle.lib()
    for s in ("orc_evolve_wave_half_flat", "orc_run", "orc_energy_total", "orc_split_run"):
        assert hasattr(L, s)


def test_params_struct_layout():
    """ctypes mirror has the C layout: 8 ints, 6 doubles, 4 ints."""
    import wavehip
    assert ctypes.sizeof(wavehip.WvParams) == 8 * 4 + 6 * 8 + 4 * 4
