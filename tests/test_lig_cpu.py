"""CPU checks of the learned-initial-guess reference (tests/lig_numpy.py) and
of the host-side pieces of the learned path.  The analytic gradient of the
NumPy reference is pinned by central finite differences; the GPU tests then
compare the HIP kernels with this reference."""
import os

import numpy as np
import pytest

import lig_numpy as R


def _rand_problem(rng, hidden=8, nlayers=2, lnet=1, B=7):
    nout = 8 * lnet * (lnet + 2)
    p = rng.normal(size=R.nparams(hidden, nlayers, nout)) * 0.3
    p[-2:] = [0.05, 0.02]
    X = rng.normal(size=(B, R.NFEAT))
    Y = rng.normal(size=(B, nout))
    base = rng.normal(size=(B, nout))
    depth = rng.random(B) * 5
    return p, X, Y, base, depth, hidden, nlayers, nout


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p, X, Y, base, depth, h, nl, no = _rand_problem(rng)
    _, g = R.loss_grad(p, X, Y, base, depth, h, nl, no)
    idx = np.concatenate([rng.choice(len(p) - 2, 60, replace=False), [len(p) - 2, len(p) - 1]])
    for k in idx:
        e = np.zeros_like(p)
        e[k] = 1e-6
        lp, _ = R.loss_grad(p + e, X, Y, base, depth, h, nl, no)
        lm, _ = R.loss_grad(p - e, X, Y, base, depth, h, nl, no)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_numpy_loss_is_zero_at_the_target():
    # NN = 0 and kappa = 0 reproduce base exactly: loss(Y = base) == 0
    rng = np.random.default_rng(4)
    p, X, Y, base, depth, h, nl, no = _rand_problem(rng)
    Ws, bs, _ = R.unpack(p, h, nl, no)
    p0 = p.copy()
    off = len(p) - 2 - (no * h + no)
    p0[off:len(p)] = 0.0  # last layer and kappa
    loss, g = R.loss_grad(p0, X, base, base, depth, h, nl, no)
    assert loss == 0.0 and np.all(g == 0.0)


def test_feature_definitions_on_a_pair():
    # two touching spheres along z: sphere 0 sees sphere 1 straight above
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 4.0]])
    F, d = R.features(pos, [2.0, 2.0], [1.31, 1.31])
    assert d[0] == pytest.approx(2.0) and d[1] == pytest.approx(6.0)   # depth below the entry surface z=-2
    assert F[0, 0] == pytest.approx(2.0 / 16) and F[1, 1] == pytest.approx(2.0 / 16)
    s0 = R.GLOBALF   # r = 4 = D lies in shell 0 [0, 1.25D)
    assert F[0, s0] == pytest.approx(1 / 20) and F[0, s0 + 3] == pytest.approx(1 / 20)   # count, u_z
    assert F[1, s0 + 3] == pytest.approx(-1 / 20)
    assert F[0, s0 + 5] == pytest.approx(2 / 20)                                         # 2uz^2 - ux^2 - uy^2
    assert F[0, s0 + 9] == pytest.approx(np.cos(4.0) / 20)                               # cos(k dz)
    assert F[0, s0 + 11] == pytest.approx(np.cos(8.0) / 20)                              # (D/r) cos(k(dz + r))
    assert F[0, 5] == pytest.approx(1 / 150)


def test_combine_and_rows_round_trip():
    # rows of b are base; combine with O = target - base, kappa = 0 gives back amnp
    rng = np.random.default_rng(5)
    order = np.array([4, 4, 3])
    offset = np.concatenate([[0], np.cumsum(2 * order * (order + 2))])
    ne = offset[-1]
    pos = rng.normal(size=(3, 3)) * 5
    a = rng.normal(size=(2, ne)) + 1j * rng.normal(size=(2, ne))
    b = rng.normal(size=(2, ne)) + 1j * rng.normal(size=(2, ne))
    ta = R.to_rows(a, pos, order, offset, 4)
    tb = R.to_rows(b, pos, order, offset, 4)
    x = R.combine(b, ta - tb, np.zeros(3), [0.0, 0.0], pos, order, offset, 4)
    np.testing.assert_allclose(x, a, atol=1e-12)


def test_parameter_count_matches_library():
    from mstm_amd import lib

    try:
        L = lib.load_library()
    except lib.MstmHipError:
        pytest.skip("library not built")
    assert L.mstm_hip_lig_nparams(128, 3, 4) == R.nparams(128, 3, 8 * 4 * 6)
    assert L.mstm_hip_lig_nparams(8, 2, 1) == R.nparams(8, 2, 24)


def test_shipped_weights_load_without_pickle():
    from mstm_amd import learned

    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the tree")
    p, hidden, nlayers, lnet, meta = learned.load_weights()
    assert (hidden, nlayers, lnet) == (128, 3, 4)
    assert len(p) == R.nparams(hidden, nlayers, 8 * lnet * (lnet + 2))
    assert np.all(np.isfinite(p))
    assert meta["configs"] >= 1 and meta["loss_final"] < meta["loss_initial"]
