"""Pins of the oracle's post-solve observables (CPU, no GPU needed).

The restatements of tranordertest, merge_to_common_origin, scatteringmatrix
and fosmexpansion are checked against the reference's own outputs for seven
clusters (tests/golden/, obs_* keys, written by tests/golden/make_golden.py
from the unmodified Fortran) and against what the physics fixes without the
reference:
  * the merged far field equals the sum of the single-sphere far fields
    with their path phases (addition theorem, independent of the merge);
  * 4 pi s00(1,1,0) is the scattering cross section of the efficiencies;
  * g = s00(1,1,1)/(3 s00(1,1,0)) equals <cos theta> of S11 integrated by
    Gauss-Legendre x trapezoid quadrature, exact for the band limit 2T;
  * the s00(1,1,w) Legendre series is the azimuthal mean of S11.
"""
import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss, legval

from conftest import golden
from oracle import oracle as O

CASES = ["cfg1_direct", "dense_l2", "fft_cyl250", "frac_agg", "sparse100", "poly30", "poly30_fft"]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def amnaddr(m, n, L):
    return (n - 1) * (L + 2) + m if m >= 0 else (-m - 1) * (L + 2) + n + 1


@pytest.mark.parametrize("name", CASES)
def test_observables_match_reference(name):
    g = golden(name)
    cl = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    T, tord = cl.translation_orders(g["obs_origin"])
    assert T == int(g["obs_T"])
    np.testing.assert_array_equal(tord, g["obs_torder"])
    a0 = cl.merge(T, tord, g["obs_origin"], g["amnp"])
    assert rel(a0, g["obs_amnp0"]) < 1e-12
    th, ct, phi = O.scattering_angles()
    sm = O.scatmat(T, g["obs_amnp0"], ct, phi)
    assert rel(sm, g["obs_scatmat"]) < 1e-12
    for key, out in zip(("s00", "s02", "sp22", "sm22"), O.fosm(T, g["obs_amnp0"])):
        assert rel(out, g["obs_" + key]) < 1e-12, key


def test_default_grid_matches_reference_loop():
    th, ct, phi = O.scattering_angles(0.0)
    assert len(th) == 361 and th[0] == -180 and th[-1] == 180
    # theta = 0 is clamped; cos of the real(4) pi is not exactly -1, so
    # theta = +-180 keep their own cosine
    assert ct[180] == 0.9999999 and ct[0] == np.cos(O.PI_R4) and ct[0] > -1.0
    assert np.all(phi[:180] == O.PI_R4) and np.all(phi[180:] == 0.0)


def _amplitude(T, a0, ct, phi):
    sm = np.zeros(16)
    sa = np.zeros(4, np.complex128)
    lib = O.lib()
    a0 = np.ascontiguousarray(a0, np.complex128)
    lib.orc_scatmat(T, a0.ctypes.data, float(ct), float(phi), 1.0, O._d(sm), sa.ctypes.data)
    return sa


def test_merge_is_the_sum_of_sphere_far_fields():
    """Far-field amplitude of the merged expansion = sum over spheres of each
    sphere's own far-field amplitude times exp(-i n.(r_i - r0))."""
    g = golden("cfg1_direct")
    cl = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    r0 = g["obs_origin"]
    T, tord = cl.translation_orders(r0, eps=1e-10)
    a0 = cl.merge(T, tord, r0, g["amnp"])
    for ct, phi in [(0.3, 0.7), (-0.8, 2.1), (0.95, -1.0)]:
        st = np.sqrt(1 - ct * ct)
        nhat = np.array([st * np.cos(phi), st * np.sin(phi), ct])
        tot = np.zeros(4, np.complex128)
        for i in range(cl.n):
            L, o = int(cl.order[i]), int(cl.offset[i])
            b = L * (L + 2)
            ai = np.zeros((2, 2, b), np.complex128)  # (pol, E/M, coef) at the sphere's own origin
            for pol in range(2):
                aL = g["amnp"][pol, o:o + b]
                aR = g["amnp"][pol, o + b:o + 2 * b]
                ai[pol, 0] = aL + aR
                ai[pol, 1] = aL - aR
            tot += _amplitude(L, ai, ct, phi) * np.exp(-1j * nhat @ (cl.pos[i] - r0))
        assert rel(_amplitude(T, a0, ct, phi), tot) < 1e-6  # translation truncation at eps 1e-10


def test_s00_energy_and_asymmetry_by_quadrature():
    g = golden("cfg1_direct")
    cl = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    sol = cl.fixedorsoln(eps=1e-14)
    qtot = cl.qtot(sol["qeff"])
    ob = O.observables(cl, sol["amnp"])
    T, s00 = ob["T"], ob["s00"]
    rv2 = float(np.sum(cl.radius ** 3)) ** (2 * float(np.float32(0.333333)))
    qsca = qtot[0, 2]
    assert abs(4.0 * s00[0, 0, 0] / (qsca * rv2) - 1.0) < 1e-5
    # g by quadrature of S11 over the sphere
    x, w = leggauss(2 * T + 2)
    nph = 4 * T + 4
    phis = 2 * np.pi * np.arange(nph) / nph
    CT, PH = np.meshgrid(x, phis, indexing="ij")
    s11 = O.scatmat(T, ob["amnp0"], CT.ravel(), PH.ravel())[:, 0].reshape(CT.shape)
    ring = s11.mean(axis=1)
    g_quad = float(np.sum(w * x * ring) / np.sum(w * ring))
    assert abs(g_quad - ob["g"]) < 1e-12
    # azimuthal mean of S11 = Legendre series of s00(1,1,w)
    assert rel(ring, legval(x, s00[:, 0, 0])) < 1e-12
