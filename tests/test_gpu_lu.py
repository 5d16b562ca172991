"""GPU direct solution (dense A = I - T H, blocked LU with the reference's
scaled partial pivoting) through the C-ABI.

* fixedorsoln with the LU solver against the reference's own LU branch
  (tests/golden/lu_cfg1, lu_poly30: 1400 and 1646 equations, orders 5-6 and
  mixed 2-6; the tail panels are partial): solution and efficiencies 1e-10.
* direct_solver with refinement steps, the factors reused
  (initialize = False), solution_error.
* At 14000 equations (200 spheres of order 5): the LU solution against the
  FP64 BiCG solve at a tight eps, and their right-hand sides under the GPU
  operator.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", ["lu_cfg1", "lu_poly30"])
def test_lu_fixedorsoln_matches_reference(gpu_ctx, name):
    from mstm_amd.solver import SOLVE_LU

    g = golden(name)
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    r = gpu_ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_LU)
    assert r["iterations"] == 0 and r["istat"] == 0
    assert rel(r["amnp"], g["amnp"]) < 1e-10, rel(r["amnp"], g["amnp"])
    assert rel(r["qeff"], g["qeff"]) < 1e-10


def test_direct_solver_refinement_and_reuse(gpu_ctx):
    g = golden("lu_cfg1")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    rng = np.random.default_rng(4)
    ne = g["amnp"].shape[1]
    p = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    x0, e0 = gpu_ctx.direct_solver(p, initialize=True, niter=0)
    assert e0 == 0.0
    x3, e3 = gpu_ctx.direct_solver(p, initialize=False, niter=3)
    assert 0.0 < e3 < 1e-14 and rel(x3, x0) < 1e-12
    # A x = p under the GPU operator: x - T H x
    thx = gpu_ctx.sphereinteraction(np.stack([x3, x3]), [False, False], [True, True])[0]
    assert rel(x3 - thx, p) < 1e-12


def test_lu_14000_equations(gpu_ctx):
    from mstm_amd import geometry as G
    from mstm_amd.solver import SOLVE_LU, SOLVE_PLAIN

    c = G.rsa_in_sphere(200, 2.0, 0.20, seed=17, ri=1.31 + 0.01j)
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    r = gpu_ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_LU)
    ne = r["amnp"].shape[1]
    assert ne == 14000
    st = gpu_ctx.stats()
    print(f"LU {ne} equations: build {st['lu_build_ms']:.1f} ms, factor {st['lu_factor_ms']:.1f} ms, "
          f"solve {st['lu_solve_ms']:.1f} ms")
    assert st["lu_factor_ms"] > 0
    x = r["amnp"]
    thx = gpu_ctx.sphereinteraction(x, [False, False], [True, True])
    t = gpu_ctx.fixedorsoln(eps=1e-20, niter=400, solver=SOLVE_PLAIN)
    assert rel(x, t["amnp"]) < 1e-7, rel(x, t["amnp"])
    b_from_x = x - thx
    b_from_t = t["amnp"] - gpu_ctx.sphereinteraction(t["amnp"], [False, False], [True, True])
    assert rel(b_from_x, b_from_t) < 1e-7


def test_lu_error_paths(gpu_ctx):
    from mstm_amd.solver import MstmHipError, SOLVE_LU

    g = golden("fft_cyl250")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    gpu_ctx.node_selection(0.0, -1, 2)
    gpu_ctx.setup_fft(0)
    with pytest.raises(MstmHipError):
        gpu_ctx.fixedorsoln(eps=1e-6, niter=100, solver=SOLVE_LU, fft_option=True)
