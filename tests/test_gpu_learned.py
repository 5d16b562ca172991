"""GPU checks of the learned initial guess: the HIP feature, sample, forward,
backward and combine kernels against the NumPy reference (tests/lig_numpy.py),
training that lowers the loss, and the learned solver (converges to the same
residual, agrees with the plain FP64 solution, and needs fewer iterations than
the mixed solver without the network on held-out clusters).

Tolerances: FP32 network arithmetic, so forward outputs 1e-4 relative to the
output scale and gradients 2e-3 relative (max-norm) to the NumPy FP64
gradient; features 1e-5 absolute (FP32 sums of O(1) terms, <= 150 terms).
"""
import os

import numpy as np
import pytest

import lig_numpy as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def trainer():
    from mstm_amd.learned import Trainer

    t = Trainer(0, hidden=128, nlayers=3, lnet=4, seed=11)
    yield t
    t.close()


def _problem(rng, B=300, nout=192):
    X = rng.normal(size=(B, 64)).astype(np.float32) * 0.5
    Y = rng.normal(size=(B, nout)).astype(np.float32)
    base = rng.normal(size=(B, nout)).astype(np.float32)
    depth = (rng.random(B) * 30).astype(np.float32)
    return X, Y, base, depth


def test_forward_matches_numpy(trainer):
    rng = np.random.default_rng(1)
    p = trainer.params.astype(np.float64)
    p[:-2] += rng.normal(size=len(p) - 2) * 0.02
    trainer.params = p.astype(np.float32)
    X, _, _, _ = _problem(rng)
    out = trainer.forward(X)
    ref, _, _ = R.forward(trainer.params, X, 128, 3, 192)
    assert rel(out, ref) < 1e-4


def test_gradient_matches_numpy(trainer):
    rng = np.random.default_rng(2)
    p = trainer.params.astype(np.float64)
    p[:-2] += rng.normal(size=len(p) - 2) * 0.02
    p[-2:] = [0.03, 0.01]
    trainer.params = p.astype(np.float32)
    X, Y, base, depth = _problem(rng)
    g, loss = trainer.grad(X, Y, base, depth)
    lref, gref = R.loss_grad(trainer.params, X, Y, base, depth, 128, 3, 192)
    assert abs(loss - lref) < 1e-4 * lref
    assert rel(g, gref) < 2e-3
    # every layer's block separately (the small ones must not hide in the max-norm)
    Ws, _, _ = R.unpack(np.arange(len(g)), 128, 3, 192)
    for W in Ws:
        sl = slice(int(W.min()), int(W.max()) + 1)
        assert rel(g[sl], gref[sl]) < 2e-3
    assert rel(g[-2:], gref[-2:]) < 2e-3


def test_training_lowers_the_loss(trainer):
    rng = np.random.default_rng(3)
    # a learnable map: Y = base + linear(X) with a phase by depth
    X, _, base, depth = _problem(rng, B=4000)
    A = rng.normal(size=(64, 192)).astype(np.float32) * 0.1
    Y = (base + X @ A).astype(np.float32)
    from mstm_amd.learned import Trainer

    t = Trainer(0, seed=5)
    t.set_data(X, Y, base, depth)
    l0 = t.loss()
    for e in range(15):
        t.epoch(256, 2e-3, seed=e)
    l1 = t.loss()
    t.close()
    assert l1 < 0.25 * l0, (l0, l1)


@pytest.fixture(scope="module")
def cluster300():
    from mstm_amd import geometry as G

    return G.rsa_in_sphere(300, 2.0, 0.30, seed=424242, ri=1.31)


def test_features_match_numpy(gpu_ctx, cluster300):
    c = cluster300
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    F, d = gpu_ctx.lig_features()
    Fr, dr = R.features(c.pos, c.radius, c.ri)
    assert np.abs(F - Fr).max() < 1e-5
    assert np.abs(d - dr).max() < 1e-4


def test_samples_and_guess_match_numpy(gpu_ctx, cluster300):
    from oracle import oracle as O

    c = cluster300
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    order, offset = gpu_ctx.orders()
    rng = np.random.default_rng(9)
    ne = gpu_ctx.number_eqns
    a = rng.normal(size=(2, ne)) + 1j * rng.normal(size=(2, ne))
    tg = np.zeros((len(order), 192), np.float32)
    tb = np.zeros((len(order), 192), np.float32)
    gpu_ctx.lig_samples_into(a, 4, tg, tb)
    np.testing.assert_allclose(tg, R.to_rows(a, c.pos, order, offset, 4), atol=2e-6 * np.abs(a).max())
    oc = O.OracleCluster(c.pos, c.radius, c.ri)
    b = np.stack([oc.mie_mult(col) for col in oc.planewave()])
    np.testing.assert_allclose(tb, R.to_rows(b, c.pos, order, offset, 4), atol=2e-6 * np.abs(b).max())
    # the starting vector from a (random) network
    from mstm_amd.learned import nparams

    p = (rng.normal(size=nparams()) * 0.05).astype(np.float32)
    p[-2:] = [0.02, 0.005]
    gpu_ctx.set_learned(p)
    x0 = gpu_ctx.lig_guess()
    Fr, dr = R.features(c.pos, c.radius, c.ri)
    O_ref, _, _ = R.forward(p, Fr.astype(np.float32), 128, 3, 192)
    xr = R.combine(b, O_ref, dr, p[-2:], c.pos, order, offset, 4)
    assert rel(x0, xr) < 1e-4
    gpu_ctx.set_learned(None)


def _weights():
    from mstm_amd import learned

    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights in the tree")
    return learned.DEFAULT_WEIGHTS


@pytest.mark.parametrize("n,seed", [(1000, 900001), (1800, 900002)])
def test_learned_solver_converges_and_saves_iterations(gpu_ctx, n, seed):
    """Held-out clusters (seeds outside the training range): the learned solve
    reaches the same residual, agrees with the plain FP64 solve to the
    accuracy that residual allows, and takes fewer FP32 iterations than the
    mixed solver started from b."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import SOLVE_LEARNED, SOLVE_MIXED, SOLVE_PLAIN

    path = _weights()
    c = G.rsa_in_sphere(n, 2.0, 0.30, seed=seed, ri=1.31)
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    gpu_ctx.load_learned(path)
    runs = {}
    for name, s in (("plain", SOLVE_PLAIN), ("mixed", SOLVE_MIXED), ("learned", SOLVE_LEARNED)):
        gpu_ctx.clear(False)
        r = gpu_ctx.fixedorsoln(eps=1e-6, niter=10000, solver=s)
        r["stats"] = gpu_ctx.stats()
        runs[name] = r
        assert r["istat"] == 0 and r["maxerr"] < 1e-6, (name, r["maxerr"])
    st = runs["learned"]["stats"]
    print(f"n={n}: iterations plain {runs['plain']['stats']['iters_fp64']:.0f} (FP64), mixed "
          f"{runs['mixed']['stats']['iters_fp32']:.0f}, learned {st['iters_fp32']:.0f} (FP32); "
          f"start residual learned {st['lig_err0']:.3e} vs b {st['lig_errb']:.3e}; fallbacks {st['lig_fallbacks']:.0f}")
    ref = runs["plain"]["amnp"]
    # residual 1e-6 in |r|^2/|b|^2 leaves ~1e-3 relative in the coefficients
    assert rel(runs["learned"]["amnp"], ref) < 2e-2
    assert rel(runs["mixed"]["amnp"], ref) < 2e-2
    assert st["lig_err0"] < st["lig_errb"]
    assert st["lig_fallbacks"] == 0
    assert st["iters_fp32"] < runs["mixed"]["stats"]["iters_fp32"]
    gpu_ctx.set_learned(None)


def test_learned_solver_tight_agreement_with_fp64(gpu_ctx):
    """Driven to |r|^2/|b|^2 < 1e-14, the learned + mixed solve and the plain
    FP64 solve of a held-out cluster agree to 1e-5 (coefficients) and 1e-6
    (efficiencies): the network only moves the start, FP64 residuals judge."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import SOLVE_LEARNED, SOLVE_PLAIN, qtotcalc

    path = _weights()
    c = G.rsa_in_sphere(1000, 2.0, 0.30, seed=900003, ri=1.31)
    gpu_ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
    gpu_ctx.load_learned(path)
    gpu_ctx.clear(False)
    rl = gpu_ctx.fixedorsoln(eps=1e-14, niter=10000, solver=SOLVE_LEARNED)
    st = gpu_ctx.stats()
    rp = gpu_ctx.fixedorsoln(eps=1e-14, niter=10000, solver=SOLVE_PLAIN)
    gpu_ctx.set_learned(None)
    assert rl["istat"] == 0 and rl["maxerr"] < 1e-14 and rp["maxerr"] < 1e-14
    assert st["lig_err0"] < 1.0  # the network's start was used
    assert rel(rl["amnp"], rp["amnp"]) < 1e-5
    assert rel(qtotcalc(c.radius, rl["qeff"]), qtotcalc(c.radius, rp["qeff"])) < 1e-6
