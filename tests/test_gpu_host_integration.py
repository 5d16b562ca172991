"""The Fortran host path: MSTM's own modules set up the cluster and call the
GPU through the ISO_C_BINDING shim (mstm_amd/fortran/mstm_hip_solver.f90);
oracle/_ref/host_integration compares each hip_* procedure with the
reference procedure it replaces on the same state.

Direct operator (20 spheres): fixedorsoln (all spheres excited, and an
excitation radius), its order-of-scattering branch (max_iterations < 0),
sphereinteraction with A and A^H columns, rhs_list and
skip_external_translation, cbicg (also with niter < 0), direct_solver before
and after a geometry change and with reused factors, all at 1e-10 relative;
the single-origin observables at 1e-12; tmatrix_solution to order 3 (the two
T-matrix files at their 11 printed digits, sphere_qeff at 1e-10).

Node-grid operator (the first 250 spheres of the reference fixture
random_in_cylinder_1000.pos): hip_node_selection with the automatic and a
specified cell size (the GPU grid must equal the host's), sphereinteraction,
fft_external_to_external_expansion and cbicg at 1e-5 (the reference keeps a
complex(4) node-kernel spectrum) with the reference's iteration count."""
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


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/host_integration not built")
def test_fortran_host_fft_through_shim(tmp_path):
    from conftest import golden

    g = golden("fft_cyl250")
    c = G.Cluster(pos=g["pos"], radius=g["radius"], ri=g["ri"])
    pos = tmp_path / "cyl250.pos"
    G.write_pos(str(pos), c)
    r = subprocess.run([BIN, str(pos), str(c.n), "fft"], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "HOST INTEGRATION OK" in r.stdout
