"""The C-ABI library loads on the CPU host and exports every symbol that
include/mstm_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os

import pytest

from mstm_amd import lib


def test_library_exists_and_exports_header_symbols():
    if not os.path.exists(lib.LIB_PATH):
        pytest.fail("libmstm_hip.so not built: run python -m mstm_amd.build")
    L = ctypes.CDLL(lib.LIB_PATH)
    syms = lib.header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_ctypes_signatures_cover_header():
    assert set(lib.header_symbols()) <= set(lib._SIGNATURES)


def test_stats_names_without_gpu():
    L = lib.load_library()
    names = L.mstm_hip_stats_names().decode().split(",")
    assert "tile_ms_fp64" in names and "solve_ms" in names


def test_create_is_clean_with_or_without_gpu():
    L = lib.load_library()
    h = ctypes.c_void_p()
    rc = L.mstm_hip_create(ctypes.byref(h), 0)
    if rc == 0:
        assert L.mstm_hip_destroy(h) == 0
    else:
        assert not h.value
