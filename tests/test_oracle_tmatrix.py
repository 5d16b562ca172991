"""Pins of the oracle's tmatrix_solution restatement (CPU, no GPU needed).

Against the reference (tests/golden/tmat5, tmat_poly6, written by
tests/golden/make_golden.py from the unmodified Fortran): the order reached,
the translation orders, the cumulative iteration counts and Q_ext/Q_abs of
every order's progress line, the T-matrix file (written by the reference
with 11 significant digits) and the per-sphere efficiency sums.

The orientation average of ranorientscatmatrix (mstm-scatprops-26.f90:1963-2530)
restated in the oracle is checked on the reference's own T-matrix files
against the reference's sm / smcf, and, without the reference, for a single
sphere: its orientation average is its fixed-orientation result, so the
(1,1), (1,4), (4,1), (4,4) coefficients equal fosmexpansion's s00.

Without the reference:
  * reciprocity, T^{pq}_{mn,kl} = (-1)^{m+k} T^{qp}_{-k l,-m n}, which ties
    the distribution of the excitation to the merge of the solution (a sign,
    index or transposition slip in either breaks it);
  * the orientation-averaged Q_ext from the T-matrix diagonal (qteff) equals
    the sum of the per-sphere extinction efficiencies.
"""
import numpy as np
import pytest

from conftest import golden
from oracle import oracle as O
from tmatrix_util import reciprocity_error

CASES = ["tmat5", "tmat_poly6"]


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", CASES)
def test_tmatrix_matches_reference(name):
    g = golden(name)
    cl = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    o = cl.tmatrix(eps=float(g["eps"]), niter=int(g["niter"]), conveps=float(g["conveps"]))
    np.testing.assert_array_equal(o["torder"], g["torder"])
    assert abs(o["vol_radius"] - float(g["vol_radius"])) < 1e-15 * float(g["vol_radius"])
    assert o["T"] == int(g["T"]) and o["istat"] == int(g["istat"])
    assert int(g["tmat_header"][0]) == o["T"]
    # progress lines: cumulative max iterations exact, Q_ext / Q_abs to the printed 5 digits
    np.testing.assert_array_equal(o["qlog"][:, 0], g["tmlog"][:, 1])
    np.testing.assert_allclose(o["qlog"][:, 1:3], g["tmlog"][:, 2:4], rtol=6e-5)
    assert o["tmat"].size == g["tmat"].size
    assert rel(o["tmat"], g["tmat"]) < 1e-9
    assert rel(o["sqeff"], g["sqeff"]) < 1e-12


def test_tmatrix_reciprocity_and_extinction():
    g = golden("tmat_poly6")
    cl = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    o = cl.tmatrix(eps=1e-14, niter=10000, conveps=1e-9)
    assert reciprocity_error(o["tmat"], o["T"]) < 1e-6
    qext = float((o["sqeff"][:, 0] * g["radius"] ** 2).sum() / o["vol_radius"] ** 2)
    assert abs(o["qteff"][0] - qext) < 1e-8 * qext
    # a planted slip: the merge taken with the wrong sign of the mode-2 excitation breaks reciprocity
    bad = o["tmat"].copy()
    bad[1::2] *= -1.0
    assert reciprocity_error(bad, o["T"]) > 1e-2


@pytest.mark.parametrize("name", CASES)
def test_ranorient_matches_reference(name):
    g = golden(name)
    sm, smcf = O.ranorient(int(g["T"]), g["tmat"])
    assert rel(sm, g["rsm"]) < 1e-13 and rel(smcf, g["rsmcf"]) < 1e-13


def test_ranorient_single_sphere_equals_fixed_orientation():
    cl = O.OracleCluster(np.zeros((1, 3)), [1.7], [1.45 + 0.03j])
    o = cl.tmatrix(eps=1e-14, conveps=0.0)
    sm, _ = O.ranorient(o["T"], o["tmat"])
    ob = O.observables(cl, cl.fixedorsoln(eps=1e-14)["amnp"])
    s00 = np.asarray(ob["s00"])  # [w][j][i] = s00(i, j, w)
    assert ob["T"] == o["T"]
    # the average reads the T-matrix into complex(4), as the reference does: 1e-8 relative here
    for i, j in ((0, 0), (0, 3), (3, 0), (3, 3)):
        assert np.abs(sm[:, j, i] - s00[:, j, i]).max() < 1e-6 * np.abs(sm[:, 0, 0]).max()
