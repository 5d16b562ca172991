"""Pins of the C oracle (CPU, no GPU needed).

Each restated function is checked against something a plausible mistake
would fail: closed forms and library routines (Wigner-d by Wigner's sum,
exact rational Clebsch-Gordan, scipy spherical Bessel functions, textbook
Lorenz-Mie), two independent formulas for the same operator (rotate-axial-
rotate vs dense vector-coupling translation matrix), a dense LU solve, and
the reference MSTM's own full-precision outputs (tests/golden/, written by
tests/golden/make_golden.py from the unmodified Fortran)."""
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

from conftest import golden
from oracle import oracle as O


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


# ---------------------------------------------------------------- special fns
def wigner_small_d(j, mp, m, beta):
    """Wigner's explicit sum for d^j_{m' m}(beta)."""
    f = math.factorial
    s = 0.0
    for k in range(0, 2 * j + 1):
        a = (j + m - k, k, mp - m + k, j - mp - k)
        if min(a) < 0:
            continue
        s += ((-1) ** (mp - m + k) * math.sqrt(f(j + mp) * f(j - mp) * f(j + m) * f(j - m)) /
              (f(a[0]) * f(a[1]) * f(a[2]) * f(a[3])) * math.cos(beta / 2) ** (2 * j + m - mp - 2 * k) *
              math.sin(beta / 2) ** (mp - m + 2 * k))
    return s


@pytest.mark.parametrize("beta", [0.0, 0.3, 1.1, 2.0, math.pi - 1e-3])
def test_rotcoef_is_wigner_d(beta):
    # rotcoef's dc(m, n(n+1)+k) is d^n_{k m}(beta) (mstm-modules-33.f90:1090-1091)
    L = 6
    dc = O.rotcoef(math.cos(beta), L, L)
    for n in range(1, L + 1):
        for m in range(-n, n + 1):
            for k in range(-n, n + 1):
                assert abs(dc[n * (n + 1) + k, m + L] - wigner_small_d(n, k, m, beta)) < 1e-13


def test_rotcoef_orthogonal():
    L = 7
    dc = O.rotcoef(0.37, L, L)
    for n in range(1, L + 1):
        D = np.array([[dc[n * (n + 1) + k, m + L] for k in range(-n, n + 1)] for m in range(-n, n + 1)])
        assert np.abs(D @ D.T - np.eye(2 * n + 1)).max() < 1e-13


def exact_cg(j1, m1, j2, m2, J):
    """Clebsch-Gordan coefficient squared with sign, from Racah's formula in
    exact rational arithmetic."""
    M = m1 + m2
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J or not abs(j1 - j2) <= J <= j1 + j2:
        return 0.0
    f = math.factorial
    pre = Fraction((2 * J + 1) * f(J + j1 - j2) * f(J - j1 + j2) * f(j1 + j2 - J), f(j1 + j2 + J + 1))
    pre *= f(J + M) * f(J - M) * f(j1 - m1) * f(j1 + m1) * f(j2 - m2) * f(j2 + m2)
    s = Fraction(0)
    for k in range(0, j1 + j2 + 1):
        d = (k, j1 + j2 - J - k, j1 - m1 - k, j2 + m2 - k, J - j2 + m1 + k, J - j1 - m2 + k)
        if min(d) < 0:
            continue
        s += Fraction((-1) ** k, math.prod(f(x) for x in d))
    sign = 1 if s >= 0 else -1
    return sign * math.sqrt(float(pre * s * s))


def test_clebsch_gordan_exact():
    L = O.lib()
    for n in range(1, 9):
        for l in range(1, 9):
            for m in (-min(n, l), -1, 0, 1, min(n, l)):
                for k in (-m, 1, -1):
                    for w in range(abs(n - l), n + l + 1):
                        ref = exact_cg(n, m, l, k, w)
                        assert abs(L.orc_cg(n, m, l, k, w) - ref) < 1e-14 * max(1.0, abs(ref)), (n, m, l, k, w)


@pytest.mark.parametrize("x", [0.3, 2.0, 7.25, 31.0])
def test_riccati_vs_scipy(x):
    n = 25
    k = np.arange(n + 1)
    psi = O.riccati_psi(n, x)
    xi = O.riccati_xi(n, x)
    ref_psi = x * spherical_jn(k, x)
    ref_chi = x * spherical_yn(k, x)
    assert np.allclose(psi, ref_psi, rtol=1e-12, atol=1e-300)
    assert np.allclose(xi.imag, ref_chi, rtol=1e-12)


def test_logderiv_complex():
    z = 3.0 + 0.2j
    d = O.logderiv(12, z)
    # D_n = psi_{n-1}/psi_n - n/z with psi_n = z j_n(z) by a high-order series check
    from mpmath import mp, besselj, sqrt as msqrt, pi as mpi  # mpmath ships with sympy-free envs too

    mp.dps = 30

    def psi(nn):
        return complex(msqrt(mpi * z / 2) * besselj(nn + 0.5, z)) * 1.0

    for nn in range(1, 12):
        ref = psi(nn - 1) / psi(nn) - nn / z
        assert abs(d[nn] - ref) < 1e-12 * abs(ref)


def bh_mie(n, x, m):
    """Bohren-Huffman a_n, b_n."""
    def psi(nn, z):
        return z * spherical_jn(nn, z)

    def dpsi(nn, z):
        return spherical_jn(nn, z) + z * spherical_jn(nn, z, derivative=True)

    def xi(nn, z):
        return z * (spherical_jn(nn, z) + 1j * spherical_yn(nn, z))

    def dxi(nn, z):
        return (spherical_jn(nn, z) + 1j * spherical_yn(nn, z)) + z * (
            spherical_jn(nn, z, derivative=True) + 1j * spherical_yn(nn, z, derivative=True))

    mx = m * x
    a = (m * psi(n, mx) * dpsi(n, x) - psi(n, x) * dpsi(n, mx)) / (m * psi(n, mx) * dxi(n, x) - xi(n, x) * dpsi(n, mx))
    b = (psi(n, mx) * dpsi(n, x) - m * psi(n, x) * dpsi(n, mx)) / (psi(n, mx) * dxi(n, x) - m * xi(n, x) * dpsi(n, mx))
    return a, b


@pytest.mark.parametrize("x,m", [(2.0, 1.5 + 0.01j), (0.5, 1.75 + 0.4j), (3.0, 1.6 + 0.02j), (2.0, 1.31 + 0j)])
def test_mie_lr_matrix_is_textbook(x, m):
    # lrtomodetran (mstm-modules-33.f90:5424-5432) maps the L/R matrix to
    # diag(-a_n, -b_n) for a non-chiral sphere
    nodr = 8
    an, qe, qs = O.mie_coef(x, m, nodr)
    qe_bh = 0.0
    for n in range(1, nodr + 1):
        A = an[n - 1]
        am = 0.5 * np.array([[A[0, 0] + A[0, 1] + A[1, 0] + A[1, 1], A[0, 0] - A[0, 1] + A[1, 0] - A[1, 1]],
                             [A[0, 0] + A[0, 1] - A[1, 0] - A[1, 1], A[0, 0] - A[0, 1] - A[1, 0] + A[1, 1]]])
        a, b = bh_mie(n, x, m)
        assert abs(am[0, 0] + a) < 1e-12 and abs(am[1, 1] + b) < 1e-12
        assert abs(am[0, 1]) < 1e-13 and abs(am[1, 0]) < 1e-13
        qe_bh += (2 * n + 1) * (a + b).real
    assert abs(qe - 2 / x ** 2 * qe_bh) < 1e-12


def test_single_sphere_cluster_reproduces_mie():
    x, m = 2.0, 1.5 + 0.01j
    oc = O.OracleCluster(np.zeros((1, 3)), [x], [m])
    r = oc.fixedorsoln()
    L = int(oc.order[0])
    _, qe, qs = O.mie_coef(x, m, L)
    assert abs(r["qeff"][0, 0, 0] - qe) < 1e-12
    assert abs(r["qeff"][0, 0, 1] - (qe - qs)) < 1e-12


@pytest.mark.parametrize("itype", [1, 3])
@pytest.mark.parametrize("no_a,no_g", [(4, 4), (3, 6), (6, 2), (2, 2)])
def test_rot_axial_equals_dense_translation(itype, no_a, no_g):
    # coefficient_translation's rot path vs gentranmatrix, both directions,
    # both shift/transpose flags (mstm-modules-33.f90:6697-6782)
    rng = np.random.default_rng(7)
    a = rng.normal(size=2 * no_a * (no_a + 2)) + 1j * rng.normal(size=2 * no_a * (no_a + 2))
    rt = np.array([1.3, -2.2, 0.7]) * (1.0 if itype == 1 else 2.5)
    for sop in (0, 1):
        for top in (0, 1):
            g1 = O.pair_translate(itype, rt, no_a, no_g, sop, top, a, force_rot=True)
            g2 = O.pair_translate(itype, rt, no_a, no_g, sop, top, a, force_rot=False) if min(no_a, no_g) < 3 else None
            if g2 is None:
                # dense path only runs below translation_switch_order: build it directly
                nodrs, nodrt = (no_g, no_a) if sop else (no_a, no_g)
                H = O.gentranmatrix(nodrs, nodrt, rt, itype)  # [p, source, target]
                aa = a.reshape(2, -1).copy()
                if sop != top:
                    aa = np.array([_mflip(no_a, v) for v in aa])
                g2 = np.stack([(H[p].T @ aa[p]) if not sop else (H[p] @ aa[p]) for p in range(2)])
                if sop != top:
                    g2 = np.array([_mflip(no_g, v) for v in g2])
                g2 = g2.ravel()
            assert rel(g1, g2) < 1e-12, (sop, top)


def _amn(m, n, L):
    return (n - 1) * (L + 2) + m if m >= 0 else (-m - 1) * (L + 2) + n + 1


def _mflip(L, v):
    out = v.copy()
    for n in range(1, L + 1):
        for m in range(1, n + 1):
            s = -1 if m % 2 else 1
            out[_amn(m, n, L)] = s * v[_amn(-m, n, L)]
            out[_amn(-m, n, L)] = s * v[_amn(m, n, L)]
    return out


def test_bicg_equals_dense_solve():
    # 4 spheres, order 3: assemble A = I - T H column by column, LU-solve,
    # and compare with cbicg driven to a tight residual
    rng = np.random.default_rng(3)
    pos = np.array([[0, 0, 0], [1.3, 0.2, 0.1], [0.1, 1.5, -0.4], [-0.9, -0.7, 1.2]], float)
    oc = O.OracleCluster(pos, [0.5] * 4, [1.6 + 0.05j] * 4)
    ne = oc.neqns
    A = np.eye(ne, dtype=complex)
    for q in range(ne):
        e = np.zeros((1, ne), complex)
        e[0, q] = 1.0
        A[:, q] -= oc.sphereinteraction(e, [0], [1])[0]
    b = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    x_lu = np.linalg.solve(A, b)
    x_cg, it, err = oc.cbicg(b, b, niter=500, eps=1e-26)
    assert rel(x_cg, x_lu) < 1e-10
    # the con_tran column is the Hermitian adjoint of T H
    y = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    ah = oc.sphereinteraction(np.stack([y]), [1], [1])[0]
    assert rel(ah, (np.eye(ne) - A).conj().T @ y) < 1e-12


# ------------------------------------------------------------ reference goldens
@pytest.mark.parametrize("case", ["cfg1_direct", "dense_l2", "frac_agg", "poly30"])
def test_oracle_matches_reference_direct(case):
    g = golden(case)
    oc = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    np.testing.assert_array_equal(oc.order, g["order"])
    assert rel(oc.mie(), g["anmie"]) < 1e-13
    a = oc.sphereinteraction(g["ain"], [0, 1], [1, 1])
    assert rel(a, g["aout"]) < 1e-12
    h = oc.external_to_external(g["ain"], [0, 1])
    assert rel(h, g["hout"]) < 1e-12
    if "amnp" in g:
        r = oc.fixedorsoln()
        assert r["iterations"] == g["iterations"]
        assert rel(r["amnp"], g["amnp"]) < 1e-12
        assert rel(r["qeff"], g["qeff"]) < 1e-12
        assert rel(oc.qtot(r["qeff"]), g["qtot"]) < 1e-12


@pytest.mark.slow
def test_oracle_matches_reference_fft():
    g = golden("fft_cyl250")
    oc = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    fi = oc.node_selection()
    np.testing.assert_array_equal(fi["cell_dim"], g["cell_dim"])
    assert abs(fi["d_cell"] - g["d_cell"]) < 1e-12 * g["d_cell"]
    assert fi["node_order"] == g["node_order"]
    a = oc.sphereinteraction(g["ain"], [0, 1], [1, 1], use_fft=True)
    # the reference stores the node kernel spectrum in complex(4)
    # (fft_translation-5.f90:31); the oracle sums node pairs in FP64
    assert rel(a, g["aout"]) < 1e-5
