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


def test_header_constants_match_python():
    """The precision and solver codes of include/mstm_hip.h are the ones the
    Python binding passes."""
    import re

    txt = open(lib.HEADER).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define MSTM_HIP_(\w+)\s+(-?\d+)", txt)}
    for name in ("FP64", "FP32", "FP64_FP32", "SOLVE_PLAIN", "SOLVE_MIXED", "SOLVE_LEARNED", "SOLVE_FGMRES",
                 "SOLVE_LU"):
        assert defs[name] == getattr(lib, name), name


def test_bench_runs_on_one_gpu_only():
    """bench.py --gpus N > 1 stops with a message instead of running."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
