"""Independent contexts driven from host threads at once (one stream, operator
and set of Krylov vectors each; the ctypes calls release the GIL): every
solve gives the bits it gives alone.  Covers the direct operator with its
graph-replayed small-cluster BiCG, the mixed solver and the FFT operator."""
import threading

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _solve(ctx, kind):
    from mstm_amd.solver import SOLVE_MIXED, SOLVE_PLAIN

    if kind == "plain":
        return ctx.fixedorsoln(eps=1e-8, niter=1000, solver=SOLVE_PLAIN)
    if kind == "mixed":
        return ctx.fixedorsoln(alpha=0.4, sinc=0.5, eps=1e-6, niter=1000, solver=SOLVE_MIXED)
    return ctx.fixedorsoln(eps=1e-6, niter=1000, solver=SOLVE_PLAIN, fft_option=True)


def test_contexts_in_threads_match_serial_solves():
    from mstm_amd.solver import FP32, FP64, HipContext

    cases = [("cfg1_direct", "plain"), ("frac_agg", "plain"), ("cfg1_direct", "mixed"), ("poly30_fft", "fft")]
    ctxs = []
    try:
        for name, kind in cases:
            g = golden(name)
            c = HipContext(0)
            c.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
            if kind == "fft":
                c.node_selection()
                c.setup_fft(FP64)
                c.setup_fft(FP32)
            ctxs.append(c)
        serial = [_solve(c, kind) for c, (_, kind) in zip(ctxs, cases)]
        for rep in range(2):
            out = [None] * len(cases)
            err = []

            def run(i):
                try:
                    out[i] = _solve(ctxs[i], cases[i][1])
                except Exception as e:  # reported below, on the main thread
                    err.append(e)

            th = [threading.Thread(target=run, args=(i,)) for i in range(len(cases))]
            for t in th:
                t.start()
            for t in th:
                t.join()
            assert not err, err
            for i, (s, o) in enumerate(zip(serial, out)):
                assert o["iterations"] == s["iterations"] and o["istat"] == s["istat"], (cases[i], rep)
                assert np.array_equal(o["amnp"], s["amnp"]), (cases[i], rep)
    finally:
        for c in ctxs:
            c.close()
