"""The Fortran host path: MSTM's own modules set up the cluster and call the
GPU through the ISO_C_BINDING shim (mstm_amd/fortran/mstm_hip_solver.f90);
oracle/_ref/host_integration compares each hip_* procedure with the
reference procedure it replaces on the same state (fixedorsoln,
sphereinteraction with A and A^H columns, cbicg) at 1e-10 relative, the
single-origin observables at 1e-12 and tmatrix_solution (to order 3: the
two T-matrix files at their 11 printed digits, sphere_qeff at 1e-10)."""
import os
import subprocess

import pytest

from mstm_amd import geometry as G

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "host_integration")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/host_integration not built")
def test_fortran_host_through_shim(tmp_path):
    c = G.rsa_in_sphere(20, 2.0, 0.10, seed=1, ri=1.5 + 0.01j)
    pos = tmp_path / "cluster.pos"
    G.write_pos(str(pos), c)
    r = subprocess.run([BIN, str(pos), str(c.n)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HOST INTEGRATION OK" in r.stdout
