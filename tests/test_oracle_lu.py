"""Pin of the oracle's direct solution (CPU, no GPU needed).

The oracle assembles A = I - T H column by column from its operator
(general_interaction_matrix with mie_mult, mstm-modules-33.f90:5854-5985) and
solves the plane-wave system exactly; the reference's own LU branch of
fixedorsoln (solution_method 'd': lu_decomposition + lu_backsubstitution,
mstm-solver-8.f90:951-1052) gives the same solution on the 20-sphere
cluster (tests/golden/lu_cfg1) to rounding.  That A is I - T H at all is
pinned by test_oracle.py::test_bicg_equals_dense_solve.
"""
import numpy as np

from conftest import golden
from oracle import oracle as O


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def test_direct_solution_matches_reference_lu():
    g = golden("lu_cfg1")
    oc = O.OracleCluster(g["pos"], g["radius"], g["ri"])
    pw = oc.planewave()
    b = np.stack([oc.mie_mult(pw[0]), oc.mie_mult(pw[1])])
    x = oc.direct_solve(b)
    assert rel(x, g["amnp"]) < 1e-12
    assert rel(oc.qeff(x, pw), g["qeff"]) < 1e-12
