"""GPU parity of the post-solve observables (translation orders, merge to the
common origin, scattering matrix, asymmetry parameter) through the C-ABI.

Against the reference's own outputs (tests/golden/, obs_* keys) on the same
solutions: orders exactly, merged coefficients, the 361 x 16 scattering
matrix, s00(1,1,0:1) and the full fosmexpansion (s00, s02, sp22, sm22 for
every w) to 1e-12 relative (FP64 throughout; only the summation order
differs).  At the 5000-sphere cluster (common-origin order about 60) against
the C oracle on the GPU's own solution.
"""
import time

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

CASES = ["cfg1_direct", "dense_l2", "fft_cyl250", "frac_agg", "sparse100", "poly30", "poly30_fft"]
TOL = 1e-12


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("case", CASES)
def test_observables_match_reference(gpu_ctx, case):
    from mstm_amd.solver import scattering_angles

    g = golden(case)
    print(f"[obs] {case}", flush=True)
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    T, tord = gpu_ctx.translation_orders(g["obs_origin"], 1e-5, 100)
    assert T == int(g["obs_T"])
    np.testing.assert_array_equal(tord, g["obs_torder"])
    a0 = gpu_ctx.merge_to_common_origin(g["amnp"])
    assert rel(a0, g["obs_amnp0"]) < TOL
    th, ct, phi = scattering_angles(0.0)
    sm = gpu_ctx.scattering_matrix(ct, phi)
    assert rel(sm, g["obs_scatmat"]) < TOL
    s00, gasym = gpu_ctx.asymmetry()
    ref = g["obs_s00"][:2, 0, 0]
    assert rel(s00, ref) < TOL
    assert abs(gasym - ref[1] / (3 * ref[0])) < TOL
    # caller-supplied common-origin coefficients give the same answers
    sm2 = gpu_ctx.scattering_matrix(ct, phi, a0=g["obs_amnp0"], T=T)
    assert rel(sm2, g["obs_scatmat"]) < TOL
    s00b, _ = gpu_ctx.asymmetry(a0=g["obs_amnp0"], T=T)
    assert rel(s00b, ref) < TOL
    # fosmexpansion: the device-resident merge and the caller's coefficients
    for fo in (gpu_ctx.fosmexpansion(), gpu_ctx.fosmexpansion(a0=g["obs_amnp0"], T=T)):
        for key, out in zip(("s00", "s02", "sp22", "sm22"), fo):
            assert rel(out, g["obs_" + key]) < TOL, key


def test_merge_of_device_solution_equals_host_solution(gpu_ctx):
    g = golden("cfg1_direct")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    r = gpu_ctx.fixedorsoln(eps=1e-6, niter=10000)
    gpu_ctx.translation_orders()
    a_dev = gpu_ctx.merge_to_common_origin(None)
    a_host = gpu_ctx.merge_to_common_origin(r["amnp"])
    np.testing.assert_array_equal(a_dev, a_host)


def test_merge_without_orders_or_solution_is_an_error(gpu_ctx):
    from mstm_amd.solver import MstmHipError

    g = golden("dense_l2")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    with pytest.raises(MstmHipError):
        gpu_ctx.merge_to_common_origin(g["amnp"])  # no translation orders yet
    gpu_ctx.translation_orders()
    with pytest.raises(MstmHipError):
        gpu_ctx.merge_to_common_origin(None)  # no solution on the device


def test_observables_5000_sphere_cluster_match_oracle(gpu_ctx):
    """The 5000-sphere x = 2 cluster: GPU observables of the GPU's own mixed
    solution against the oracle's restatement on the same solution."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import FP32, FP64, SOLVE_MIXED, scattering_angles
    from oracle import oracle as O

    cl = G.rsa_in_sphere(5000, 2.0, 0.30, seed=5, ri=1.31 + 0.0j)
    gpu_ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    gpu_ctx.setup_direct(FP64)
    gpu_ctx.setup_direct(FP32)
    r = gpu_ctx.fixedorsoln(eps=1e-6, niter=2000, solver=SOLVE_MIXED)
    assert r["istat"] == 0
    t0 = time.perf_counter()
    ob = gpu_ctx.observables(None, copy=True)
    t_gpu = time.perf_counter() - t0
    oc = O.OracleCluster(cl.pos, cl.radius, cl.ri)
    T, tord = oc.translation_orders(None, 1e-5, 100)
    print(f"[obs] 5000 spheres: T = {ob['T']} (oracle {T}), GPU observables {t_gpu:.3f} s", flush=True)
    assert ob["T"] == T
    np.testing.assert_array_equal(ob["torder"], tord)
    a0 = oc.merge(T, tord, cl.pos.mean(axis=0), r["amnp"])
    assert rel(ob["amnp0"], a0) < 1e-11
    th, ct, phi = scattering_angles(0.0)
    sel = np.arange(0, 361, 15)
    sm = O.scatmat(T, a0, ct[sel], phi[sel])
    assert rel(ob["scatmat"][sel], sm) < 1e-11
    # g from the s00 moments against <cos theta> of S11 integrated over the sphere
    x, w = np.polynomial.legendre.leggauss(2 * T + 2)
    nph = 4 * T + 4
    CT, PH = np.meshgrid(x, 2 * np.pi * np.arange(nph) / nph, indexing="ij")
    ring = gpu_ctx.scattering_matrix(CT.ravel(), PH.ravel())[:, 0].reshape(CT.shape).mean(axis=1)
    g_quad = float(np.sum(w * x * ring) / np.sum(w * ring))
    assert abs(ob["g"] - g_quad) < 1e-10
    # the full generalized-spherical expansion against the oracle's (T about 60)
    for key, got, want in zip(("s00", "s02", "sp22", "sm22"), gpu_ctx.fosmexpansion(), O.fosm(T, a0)):
        assert rel(got, want) < 1e-10, key


def test_fosmexpansion_order_mismatch_is_an_error(gpu_ctx):
    from mstm_amd.solver import MstmHipError

    g = golden("cfg1_direct")
    gpu_ctx.set_spheres(g["pos"], g["radius"], g["ri"], 1.0, 1e-6)
    T, _ = gpu_ctx.translation_orders(g["obs_origin"], 1e-5, 100)
    gpu_ctx.merge_to_common_origin(g["amnp"])
    W = 2 * T + 3
    outs = [np.zeros((W, 4, 4)) for _ in range(4)]
    from mstm_amd.solver import _d

    rc = gpu_ctx.L.mstm_hip_fosmexpansion(gpu_ctx.h, T + 1, None, *[_d(o) for o in outs])
    assert rc != 0 and not any(np.any(o) for o in outs)


def test_sphere_at_the_common_origin(gpu_ctx):
    """A symmetric cluster whose centre sphere sits exactly at the mean
    position: its expansion is carried over by the zero-translation copy
    (mtransfer) in both the GPU merge and the oracle."""
    from mstm_amd.solver import scattering_angles
    from oracle import oracle as O

    d = 4.5
    pos = np.array([[0.0, 0.0, 0.0], [d, 0, 0], [-d, 0, 0], [0, d, 0], [0, -d, 0], [0, 0, d], [0, 0, -d]])
    radius = np.full(len(pos), 1.5)
    ri = np.full(len(pos), 1.45 + 0.02j)
    gpu_ctx.set_spheres(pos, radius, ri, 1.0, 1e-6)
    r = gpu_ctx.fixedorsoln(eps=1e-10, niter=10000)
    T, tord = gpu_ctx.translation_orders()
    oc = O.OracleCluster(pos, radius, ri)
    To, to = oc.translation_orders(np.zeros(3), 1e-5, 100)
    assert T == To and np.array_equal(tord, to) and tord[0] == gpu_ctx.sphere_order[0]
    a0 = gpu_ctx.merge_to_common_origin(None)
    ref = oc.merge(T, to, np.zeros(3), r["amnp"])
    assert rel(a0, ref) < TOL
    th, ct, phi = scattering_angles(0.0)
    assert rel(gpu_ctx.scattering_matrix(ct, phi), O.scatmat(T, ref, ct, phi)) < TOL
