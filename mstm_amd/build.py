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
SOURCES = ["context", "tables", "direct", "mie", "krylov", "fft", "learned", "batch", "observe", "lu"]
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


def _stale(name: str) -> bool:
    out = os.path.join(OBJ, name + ".o")
    if not os.path.exists(out):
        return True
    deps = [os.path.join(CSRC, name + ".hip"), os.path.join(os.path.dirname(HERE), "include", "mstm_hip.h")]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(out)


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every translation unit (only the out-of-date ones if
    incremental) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    names = [n for n in SOURCES if os.path.exists(os.path.join(CSRC, n + ".hip"))]
    todo = [n for n in names if _stale(n)] if incremental else names
    if todo:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, n + ".o") for n in names]
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
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
