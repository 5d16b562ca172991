"""GPU checks of the boundary's optional branches, through the C-ABI:

* fixedorsoln / cbicg with max_iterations < 0 (the order-of-scattering
  series of cbicg, mstm-solver-8.f90:785-825), = 0 (no solve) and
  excited_spheres (mstm-scatprops-26.f90:217-284) against the reference's
  own outputs (tests/golden/boundary_cfg1.npz) at the FP64 bar 1e-10;
* sphereinteraction's rhs_list and skip_external_translation;
* node_selection with a specified cell size and a target box against the
  reference's grids, and the FFT product on such a grid against the oracle
  (1e-5: the node-kernel spectrum is kept in FP32, as the reference's
  complex(4));
* state rules of the context: new Mie data after an LU solve refactor
  (compared with a fresh context), and the mixed solver closes a system as
  soon as its iteration budget is spent;
* the learned start is judged in FP64: with a loose eps its FP32 start
  residual is below eps, and the answer's reported error must still be the
  FP64 residual of the returned solution (an FP32 residual differs from it
  at the 1e-8 level).
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-10


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def excited_mask(pos, frac):
    d = np.sqrt(((pos - pos.mean(axis=0)) ** 2).sum(axis=1))
    return d <= frac * d.max()


@pytest.fixture
def bcase(gpu_ctx):
    g = golden("boundary_cfg1")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    return g


@pytest.mark.parametrize("solver", [0, 1, 3])
@pytest.mark.parametrize("tag,niter,eps", [("oos", -40, 1e-8), ("oos_short", -5, 1e-8), ("noiter", 0, 1e-6)])
def test_fixedorsoln_optional_branches_match_reference(gpu_ctx, bcase, tag, niter, eps, solver):
    # every iterative solver letter takes the reference's branch for niter <= 0
    g = bcase
    r = gpu_ctx.fixedorsoln(eps=eps, niter=niter, solver=solver)
    assert r["iterations"] == int(g[f"{tag}_iterations"]) and r["istat"] == int(g[f"{tag}_istat"])
    assert r["maxerr"] == pytest.approx(float(g[f"{tag}_maxerr"]), rel=1e-8, abs=1e-300)
    assert rel(r["amnp"], g[f"{tag}_amnp"]) < FP64_TOL
    assert rel(r["qeff"], g[f"{tag}_qeff"]) < FP64_TOL


def test_fixedorsoln_excited_spheres_match_reference(gpu_ctx, bcase):
    g = bcase
    ex = excited_mask(g["pos"], 0.8)
    assert ex.sum() == int(g["excited_count"])
    r = gpu_ctx.fixedorsoln(alpha=0.3, sinc=0.4, eps=1e-6, niter=10000, excited_spheres=ex)
    assert r["iterations"] == int(g["excited_iterations"]) and r["istat"] == int(g["excited_istat"])
    assert rel(r["amnp"], g["excited_amnp"]) < FP64_TOL
    assert rel(r["qeff"], g["excited_qeff"]) < FP64_TOL
    # all flags set is the full excitation
    r1 = gpu_ctx.fixedorsoln(alpha=0.3, sinc=0.4, excited_spheres=np.ones(len(ex), bool))
    r2 = gpu_ctx.fixedorsoln(alpha=0.3, sinc=0.4)
    assert np.array_equal(r1["amnp"], r2["amnp"])


def test_cbicg_order_of_scattering_and_zero_iterations(gpu_ctx, bcase):
    from oracle import oracle as O

    g = bcase
    oc = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    rng = np.random.default_rng(4)
    b = oc.mie_mult(oc.planewave()[1])
    x0 = rng.normal(size=oc.neqns) + 1j * rng.normal(size=oc.neqns)
    for niter, eps in ((-7, 1e-30), (-40, 1e-8)):
        xo, ito, eo = oc.cbicg(b, x0, niter=niter, eps=eps)
        xg, itg, eg = gpu_ctx.cbicg(niter, eps, b, x0)
        assert itg == ito and rel(xg, xo) < FP64_TOL
        assert eg == pytest.approx(eo, rel=1e-8, abs=1e-300)
    xg, itg, eg = gpu_ctx.cbicg(0, 1e-6, b, x0)
    assert itg == 0 and eg == 0.0 and not np.any(xg)
    xg, itg, eg = gpu_ctx.cbicg(-5, 1e-6, np.zeros_like(b), x0)
    assert itg == 0 and np.array_equal(xg, x0)


def test_sphereinteraction_rhs_list_and_skip(gpu_ctx, bcase):
    rng = np.random.default_rng(5)
    ne = gpu_ctx.number_eqns
    ain = rng.normal(size=(3, ne)) + 1j * rng.normal(size=(3, ne))
    full = gpu_ctx.sphereinteraction(ain, [False, True, False])
    part = gpu_ctx.sphereinteraction(ain, [False, True, False], rhs_list=[True, False, True])
    assert np.array_equal(part[0], full[0]) and np.array_equal(part[2], full[2]) and not np.any(part[1])
    skip = gpu_ctx.sphereinteraction(ain, [False, True, False], skip_external_translation=True)
    assert not np.any(skip)


@pytest.mark.parametrize("tag", ["dspec", "box", "box_dspec"])
def test_node_selection_options_match_reference(gpu_ctx, bcase, tag):
    from oracle import oracle as O

    g = bcase
    pad = float(g["grid_box_pad"]) if "box" in tag else 0.0
    kw = dict(target_min=g["pos"].min(0) - pad, target_max=g["pos"].max(0) + pad) if pad else {}
    if "dspec" in tag:
        kw.update(d_specified=True, d_cell=float(g["grid_dspec_in"]))
    info = gpu_ctx.node_selection(0.0, -1, 2, **kw)
    np.testing.assert_array_equal(info["cell_dim"], g[f"grid_{tag}_cell_dim"])
    assert abs(info["d_cell"] - float(g[f"grid_{tag}_d_cell"])) < 1e-12 * float(g[f"grid_{tag}_d_cell"])
    assert info["node_order"] == int(g[f"grid_{tag}_node_order"])
    assert info["fva"] == pytest.approx(float(g[f"grid_{tag}_fva"]), rel=1e-12)
    # the grid node_selection keeps: box minimum and cell of each sphere (:740-752)
    tmin = kw.get("target_min", g["pos"].min(0))
    np.testing.assert_allclose(info["cell_boundary"], tmin, rtol=0, atol=0)
    sn = np.floor((g["pos"] - tmin) / info["d_cell"]).astype(int) + 1
    sn = np.clip(sn, 1, info["cell_dim"])
    np.testing.assert_array_equal(info["sphere_node"], sn)
    assert info["number_neighbor_nodes"] == 18
    # the FFT product on that grid against the oracle's on the same grid
    oc = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    oc.node_selection(0.0, -1, 2, **kw)
    gpu_ctx.setup_fft(0)
    rng = np.random.default_rng(6)
    ain = rng.normal(size=(2, oc.neqns)) + 1j * rng.normal(size=(2, oc.neqns))
    out = gpu_ctx.sphereinteraction(ain, [False, True], fft_option=True)
    ref = oc.sphereinteraction(ain, [0, 1], [1, 1], use_fft=True)
    assert rel(out, ref) < 1e-5


def test_node_selection_rejects_what_is_not_on_the_path(gpu_ctx, bcase):
    from mstm_amd.solver import MstmHipError

    with pytest.raises(MstmHipError, match="local_host"):
        gpu_ctx.node_selection(0.0, -1, 2, local_host=3)
    with pytest.raises(MstmHipError, match="no extent"):
        gpu_ctx.node_selection(0.0, -1, 2, target_min=[0, 0, 0], target_max=[10, 10, 0])


def test_set_mie_after_lu_refactors(gpu_ctx):
    """New Mie data (other orders) after an LU solve: the next LU solve uses
    the new matrix, as a fresh context does (the factors are dropped with the
    rest of the state built from the old orders)."""
    from mstm_amd.solver import SOLVE_LU, HipContext

    g = golden("cfg1_direct")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    r1 = gpu_ctx.fixedorsoln(solver=SOLVE_LU)
    fresh = HipContext(0)
    try:
        fresh.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-2)  # fewer orders
        order, an = fresh.sphere_order.copy(), fresh.an_mie()
        assert order.max() < gpu_ctx.sphere_order.max()
        rf = fresh.fixedorsoln(solver=SOLVE_LU)
    finally:
        fresh.close()
    gpu_ctx.set_mie(order, an)
    r2 = gpu_ctx.fixedorsoln(solver=SOLVE_LU)
    assert r2["amnp"].shape == rf["amnp"].shape != r1["amnp"].shape
    assert rel(r2["amnp"], rf["amnp"]) < 1e-12
    # the merge then refuses to read a solution of the old layout
    gpu_ctx.set_mie(order, an)
    gpu_ctx.translation_orders()
    with pytest.raises(Exception):
        gpu_ctx.merge_to_common_origin(None)


def test_mixed_solver_stops_when_the_budget_is_spent(gpu_ctx):
    from mstm_amd.solver import SOLVE_MIXED

    g = golden("poly30")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    gpu_ctx.clear(False)
    r = gpu_ctx.fixedorsoln(eps=1e-12, niter=4, solver=SOLVE_MIXED)
    st = gpu_ctx.stats()
    assert r["istat"] == 1 and r["iterations"] == 4
    # one inner solve of 4 steps per polarization, then one FP64 residual each
    assert st["iters_fp32"] == 8 and st["refinements"] == 1


def test_learned_start_is_judged_in_fp64(gpu_ctx):
    import os

    from mstm_amd import learned
    from mstm_amd.geometry import rsa_in_sphere
    from mstm_amd.solver import SOLVE_LEARNED
    from oracle import oracle as O

    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the tree")
    c = rsa_in_sphere(300, 2.0, 0.30, seed=900011, ri=1.31)
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    gpu_ctx.load_learned()
    gpu_ctx.clear(False)
    r = gpu_ctx.fixedorsoln(eps=0.9, niter=10000, solver=SOLVE_LEARNED)
    st = gpu_ctx.stats()
    gpu_ctx.set_learned(None)
    assert st["lig_err0"] < 0.9 and r["iterations"] == 0  # the start is accepted without an inner solve
    # |r|^2/|b|^2 of the returned second polarization with the FP64 operator
    x = r["amnp"][1:2]
    oc = O.OracleCluster(c.pos, c.radius, c.ri)
    bb = oc.mie_mult(oc.planewave()[1])
    thx = gpu_ctx.sphereinteraction(x, [False])[0]
    res = bb - x[0] + thx
    e64 = float(np.vdot(res, res).real / np.vdot(bb, bb).real)
    assert r["maxerr"] == pytest.approx(e64, rel=1e-10)
    assert abs(st["lig_err0"] - e64) > 1e-12 * e64  # the FP32 start residual is another number
