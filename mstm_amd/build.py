"""Build libmstm_hip.so (gfx950) in-tree: mstm_amd/libmstm_hip.so.

Each .hip translation unit is compiled in parallel with hipcc, then linked
into one shared library.  Run as `python -m mstm_amd.build` or through
__graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmstm_hip.so")
SOURCES = ["context", "tables", "direct", "mie", "krylov", "fft", "learned"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MSTM_OFFLOAD_ARCH", "gfx950")
# packed FP32 ops stay disabled for all device code (known-bad forms on this pool)
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _compile(name: str) -> str:
    src = os.path.join(CSRC, name + ".hip")
    out = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{r.stderr[-4000:]}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    names = [n for n in SOURCES if os.path.exists(os.path.join(CSRC, n + ".hip"))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(names))) as ex:
        objs = list(ex.map(_compile, names))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
