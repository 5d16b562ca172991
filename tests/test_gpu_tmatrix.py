"""GPU tmatrix_solution (mstm-solver-8.f90:22-332) through the C-ABI.

* Reference parity (tests/golden/tmat5, tmat_poly6): the solver's FP64
  iterates follow cbicg's, so the order reached, every progress line's
  cumulative iteration count, the T-matrix (the reference file holds 11
  significant digits: 1e-9) and the per-sphere efficiency sums agree; the
  orientation-averaged scattering-matrix expansion of the GPU T-matrix
  (ranorientscatmatrix, restated in the oracle) matches the reference's.
* Oracle parity at a tight solution eps (1e-10 relative), a sphere exactly at
  the common origin (the zero-translation copy of the distribution), and the
  mixed solver against the FP64 one.
* At a size the oracle does not reach in seconds (60 spheres): reciprocity of
  the T-matrix and the T-matrix extinction against the per-sphere sum.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import golden
from tmatrix_util import reciprocity_error, tmat_size

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", ["tmat5", "tmat_poly6"])
def test_tmatrix_matches_reference(gpu_ctx, name):
    g = golden(name)
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    r = gpu_ctx.tmatrix_solution(eps=float(g["eps"]), niter=int(g["niter"]), conveps=float(g["conveps"]))
    assert r["T_in"] == int(g["T_in"]) and r["T"] == int(g["T"]) and r["istat"] == int(g["istat"])
    np.testing.assert_array_equal(r["qlog"][:, 0], g["tmlog"][:, 1])
    np.testing.assert_allclose(r["qlog"][:, 1:3], g["tmlog"][:, 2:4], rtol=6e-5)
    assert r["tmat"].size == g["tmat"].size
    assert rel(r["tmat"], g["tmat"]) < 1e-9, rel(r["tmat"], g["tmat"])
    assert rel(r["sqeff"], g["sqeff"]) < 1e-10
    # random orientation: the orientation average of the GPU T-matrix against the
    # reference's average of its own (oracle restatement of ranorientscatmatrix)
    from oracle import oracle as O

    sm, smcf = O.ranorient(r["T"], r["tmat"])
    assert rel(sm, g["rsm"]) < 1e-8 and rel(smcf, g["rsmcf"]) < 1e-8


def _oracle_tmatrix(pos, radius, ri, eps, conveps):
    from oracle import oracle as O

    return O.OracleCluster(pos, radius, ri).tmatrix(eps=eps, niter=10000, conveps=conveps)


def test_tmatrix_matches_oracle_tight(gpu_ctx):
    g = golden("tmat_poly6")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    r = gpu_ctx.tmatrix_solution(eps=1e-14, niter=10000, conveps=1e-9)
    o = _oracle_tmatrix(g["pos"], g["radius"], g["ri"], 1e-14, 1e-9)
    assert r["T"] == o["T"] and r["maxiter"] == o["maxiter"]
    assert rel(r["tmat"], o["tmat"]) < 1e-10
    assert rel(r["sqeff"], o["sqeff"]) < 1e-10
    assert rel(r["qteff"], o["qteff"]) < 1e-10
    # mixed FP32/FP64 solver: same T-matrix to the solution accuracy
    m = gpu_ctx.tmatrix_solution(eps=1e-12, niter=10000, conveps=1e-9, solver=1)
    assert m["T"] == o["T"] and m["istat"] == 0
    assert rel(m["tmat"], o["tmat"]) < 1e-5


def test_tmatrix_sphere_at_origin(gpu_ctx):
    pos = np.array([[-2.5, 0.3, 0.0], [0.0, 0.0, 0.0], [2.5, -0.3, 0.0]])
    radius = np.array([1.0, 0.8, 1.1])
    ri = np.full(3, 1.5 + 0.02j)
    gpu_ctx.set_spheres(pos, radius, ri, 1.0, 1e-6)
    r = gpu_ctx.tmatrix_solution(eps=1e-10, niter=10000, conveps=1e-6)
    o = _oracle_tmatrix(pos, radius, ri, 1e-10, 1e-6)
    assert r["T"] == o["T"] and r["maxiter"] == o["maxiter"]
    assert rel(r["tmat"], o["tmat"]) < 1e-10


def test_tmatrix_reciprocity_60_spheres(gpu_ctx):
    from mstm_amd import geometry as G

    c = G.rsa_in_sphere(60, 1.0, 0.15, seed=31, ri=1.5 + 0.01j)
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    r = gpu_ctx.tmatrix_solution(eps=1e-12, niter=10000, conveps=1e-7)
    assert r["istat"] == 0 and r["T"] >= 8
    assert reciprocity_error(r["tmat"], r["T"]) < 1e-5
    vr = float(np.sum(c.radius ** 3)) ** float(np.float32(0.333333))
    qext = float((r["sqeff"][:, 0] * c.radius ** 2).sum() / vr ** 2)
    assert abs(r["qteff"][0] - qext) < 1e-6 * qext


MSTM_HIP_EARG, MSTM_HIP_EUNSUP = 1, 4  # include/mstm_hip.h


def test_tmatrix_error_paths(gpu_ctx):
    g = golden("tmat5")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    T, _ = gpu_ctx.translation_orders()
    L = gpu_ctx.L
    tm = np.zeros(tmat_size(T), np.complex128)
    o = np.zeros(4)
    i = np.zeros(4, np.int32)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    args = (dp(tm.view(np.float64)),)
    rc = L.mstm_hip_tmatrix_solution(gpu_ctx.h, 1e-6, 100, 1e-6, 0, *args, tm.size - 1, None, None, None, ip(i),
                                     ip(i[1:]), dp(o), ip(i[2:]))
    assert rc == MSTM_HIP_EARG
    rc = L.mstm_hip_tmatrix_solution(gpu_ctx.h, 1e-6, 100, 1e-6, 2, *args, tm.size, None, None, None, ip(i),
                                     ip(i[1:]), dp(o), ip(i[2:]))
    assert rc == MSTM_HIP_EUNSUP
