"""Pins of the oracle's optional branches of the solver and grid entry
points (CPU): fixedorsoln with max_iterations < 0 (cbicg's order-of-scattering
series, mstm-solver-8.f90:785-825), max_iterations = 0 (:513-522) and
excited_spheres (sphereplanewavecoef, mstm-scatprops-26.f90:217-284), and
node_selection with a specified cell size and a target box
(fft_translation-5.f90:654-731).

Against the reference: tests/golden/boundary_cfg1.npz (written by
tests/golden/make_golden.py through the unmodified Fortran).  Without it:
the series is the partial sum of (T H)^k b of the dense operator, its limit
is the LU solution, and an excited-sphere right-hand side is the full one
with the other spheres' blocks zeroed."""
import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def case():
    g = golden("boundary_cfg1")
    return g, O.OracleCluster(g["pos"], g["radius"], g["ri"])


def excited_mask(pos, frac):
    d = np.sqrt(((pos - pos.mean(axis=0)) ** 2).sum(axis=1))
    return d <= frac * d.max()


@pytest.mark.parametrize("tag,niter,eps", [("oos", -40, 1e-8), ("oos_short", -5, 1e-8), ("noiter", 0, 1e-6)])
def test_order_of_scattering_and_no_solve_match_reference(case, tag, niter, eps):
    g, oc = case
    r = oc.fixedorsoln(niter=niter, eps=eps)
    assert r["iterations"] == int(g[f"{tag}_iterations"])
    assert r["istat"] == int(g[f"{tag}_istat"])
    assert r["maxerr"] == pytest.approx(float(g[f"{tag}_maxerr"]), rel=1e-9, abs=1e-300)
    assert rel(r["amnp"], g[f"{tag}_amnp"]) < 1e-12
    assert rel(r["qeff"], g[f"{tag}_qeff"]) < 1e-12


def test_excited_spheres_match_reference(case):
    g, oc = case
    ex = excited_mask(g["pos"], 0.8)
    assert ex.sum() == int(g["excited_count"]) and 0 < ex.sum() < len(ex)
    r = oc.fixedorsoln(alpha=0.3, sinc=0.4, excited=ex.astype(np.int32))
    assert r["iterations"] == int(g["excited_iterations"]) and r["istat"] == int(g["excited_istat"])
    assert rel(r["amnp"], g["excited_amnp"]) < 1e-12
    assert rel(r["qeff"], g["excited_qeff"]) < 1e-12


def test_excited_planewave_is_masked_full_planewave(case):
    g, oc = case
    ex = excited_mask(g["pos"], 0.8).astype(np.int32)
    full = oc.planewave(0.3, 0.4)
    part = oc.planewave(0.3, 0.4, excited=ex)
    keep = np.zeros(oc.neqns, bool)
    for i in range(oc.n):
        keep[oc.offset[i]:oc.offset[i + 1]] = bool(ex[i])
    assert np.array_equal(part[:, keep], full[:, keep]) and not np.any(part[:, ~keep])


def test_order_of_scattering_is_the_neumann_series(case):
    """x_K = sum_{k=0..K} (T H)^k b with T H = I - A of the dense operator;
    run out after K terms (iterations K + 1, maxerr 0); converged, it nears
    the LU solution A^{-1} b to within the last term."""
    g, oc = case
    A = oc.interaction_matrix()
    M = np.eye(oc.neqns) - A
    b = np.array([oc.mie_mult(p) for p in oc.planewave()])
    for p in range(2):
        x, t = b[p].copy(), b[p].copy()
        for _ in range(5):
            t = M @ t
            x = x + t
        xo, it, err = oc.cbicg(b[p], b[p], niter=-5, eps=1e-30)
        assert it == 6 and err == 0.0
        assert rel(xo, x) < 1e-12
    xs = oc.direct_solve(b)
    r = oc.fixedorsoln(niter=-40, eps=1e-8)
    assert rel(r["amnp"], xs) < 1e-4
    assert rel(r["amnp"], xs) > 1e-7  # a truncated series, not the exact solution


def test_cbicg_niter_zero_and_zero_rhs(case):
    g, oc = case
    b = oc.mie_mult(oc.planewave()[0])
    x, it, err = oc.cbicg(b, b, niter=0)
    assert it == 0 and err == 0.0 and not np.any(x)
    x, it, err = oc.cbicg(np.zeros_like(b), b, niter=-5)
    assert it == 0 and np.array_equal(x, b)  # pnp = 0 returns at once


@pytest.mark.parametrize("tag", ["dspec", "box", "box_dspec"])
def test_node_selection_options_match_reference(case, tag):
    g, oc = case
    pad = float(g["grid_box_pad"]) if "box" in tag else 0.0
    kw = dict(target_min=g["pos"].min(0) - pad, target_max=g["pos"].max(0) + pad) if pad else {}
    if "dspec" in tag:
        kw.update(d_specified=True, d_cell=float(g["grid_dspec_in"]))
    f = oc.node_selection(0.0, -1, 2, **kw)
    np.testing.assert_array_equal(f["cell_dim"], g[f"grid_{tag}_cell_dim"])
    assert f["d_cell"] == pytest.approx(float(g[f"grid_{tag}_d_cell"]), rel=1e-14)
    assert f["node_order"] == int(g[f"grid_{tag}_node_order"])
    assert f["fva"] == pytest.approx(float(g[f"grid_{tag}_fva"]), rel=1e-14)


def test_node_selection_specified_d_closed_form(case):
    g, oc = case
    ext = g["pos"].max(0) - g["pos"].min(0)
    d0 = 3.3
    f = oc.node_selection(0.0, 2, 2, d_specified=True, d_cell=d0)

    def n235(n):
        while True:
            m = n
            for p in (2, 3, 5):
                while m % p == 0:
                    m //= p
            if m == 1:
                return n
            n += 1

    cd = [n235(int(np.ceil(e / d0))) for e in ext]
    np.testing.assert_array_equal(f["cell_dim"], cd)
    assert f["d_cell"] == pytest.approx(max(ext / np.array(cd)), rel=1e-15)
    assert f["node_order"] == 2
