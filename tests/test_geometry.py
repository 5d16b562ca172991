"""Pins of the host-side geometry inputs (CPU): the seeded generators that
make every workload's cluster, and the position-file reader.

The reference's own random generator is unseeded
(random_configuration-10.f90:77-80), so the workloads come from
mstm_amd.geometry; a regression there would silently change every
workload, hence these checks of what each generator promises:
non-overlap, the bounding region, the volume fraction, determinism per
seed, touching neighbours and the fractal law of the aggregate, and the
spacing of the Fibonacci directions.

The reader follows read_sphere_data_input_file (mstm-input-37.f90:
2871-2924).  mstm-input-37.f90 does not compile unmodified here (amdflang
rejects its duplicate DATA initialisations of run_file and
random_configuration_output_file, :78-80), so the reader is pinned by the
contract its code states: numbers per line -> radius from 4, index from 6,
second index from 8, defaults 1 and 1+0i, list-directed complex as two
numbers or (re,im), then length_scale_factor and ref_index_scale_factor."""
import math
import os

import numpy as np
import pytest

from mstm_amd import geometry as G

REF_EXAMPLES = "/root/reference/examples"


def pair_min_gap(pos, a):
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    d[np.arange(len(pos)), np.arange(len(pos))] = np.inf
    return float(d.min() - 2 * a)


# ------------------------------------------------------------ generators
@pytest.mark.parametrize("n,a,fv,seed", [(20, 2.0, 0.10, 1), (300, 1.0, 0.25, 4), (500, 0.7, 0.30, 9)])
def test_rsa_in_sphere(n, a, fv, seed):
    c = G.rsa_in_sphere(n, a, fv, seed=seed, ri=1.31)
    R = a * (n / fv) ** (1.0 / 3.0)
    assert c.n == n and np.all(c.radius == a) and np.all(c.ri == 1.31)
    assert pair_min_gap(c.pos, a) >= -1e-12  # touching allowed, no overlap
    r = np.linalg.norm(c.pos, axis=1)
    # centres inside R - a (these cases complete without the jam fallback)
    # and filling it: the outermost centre lies close to the bound
    assert R - a - 0.1 * a < r.max() <= R - a + 1e-12
    assert n * a ** 3 / R ** 3 == pytest.approx(fv, rel=1e-12)
    # same seed, same cluster; another seed, another one
    assert np.array_equal(G.rsa_in_sphere(n, a, fv, seed=seed, ri=1.31).pos, c.pos)
    assert not np.array_equal(G.rsa_in_sphere(n, a, fv, seed=seed + 1, ri=1.31).pos, c.pos)


def test_rsa_in_sphere_workload_radius():
    # the 5000-sphere workload: R = a (N / fv)^(1/3) = 51.087 at a = 2, fv = 0.30
    assert 2.0 * (5000 / 0.30) ** (1 / 3) == pytest.approx(51.087, abs=1e-3)


def test_rsa_in_cylinder():
    n, a, fv = 400, 1.0, 0.25
    c = G.rsa_in_cylinder(n, a, fv, seed=2, ri=1.6 + 0.02j)
    Rc = a * (n / (3 * fv)) ** (1 / 3)
    L = 4 * Rc  # length = 2 diameters of the cylinder
    assert n * (4 / 3) * math.pi * a ** 3 / (math.pi * Rc ** 2 * L) == pytest.approx(fv, rel=1e-12)
    assert pair_min_gap(c.pos, a) >= -1e-12
    assert np.sqrt(c.pos[:, 0] ** 2 + c.pos[:, 1] ** 2).max() <= Rc - a + 1e-12
    assert np.abs(c.pos[:, 2]).max() <= L / 2 - a + 1e-12
    # the 1000-sphere x = 3 workload: Rc = 33.02, L = 132.08
    assert 3.0 * (1000 / 0.75) ** (1 / 3) == pytest.approx(33.02, abs=5e-3)


def fractal_fit(pos, a, kmin=30):
    """Least-squares (Df, kf) of N = kf (Rg/a)^Df over the growing aggregate
    (first k monomers, k = kmin..N)."""
    ks, rg = [], []
    for k in range(kmin, len(pos) + 1, 5):
        P = pos[:k]
        rg.append(math.sqrt(((P - P.mean(0)) ** 2).sum(1).mean()))
        ks.append(k)
    A = np.stack([np.log(np.array(rg) / a), np.ones(len(ks))], 1)
    df, lkf = np.linalg.lstsq(A, np.log(ks), rcond=None)[0]
    return df, math.exp(lkf)


def test_fractal_aggregate():
    n, a = 400, 0.5
    c = G.fractal_aggregate(n, a, 1.8, 1.3, seed=3, ri=1.75 + 0.4j)
    assert pair_min_gap(c.pos, a) >= -1e-9
    d = np.sqrt(((c.pos[:, None, :] - c.pos[None, :, :]) ** 2).sum(-1))
    d[np.arange(n), np.arange(n)] = np.inf
    # every monomer touches at least one other
    assert np.abs(d.min(axis=1) - 2 * a).max() < 1e-9
    assert np.abs(c.pos.mean(0)).max() < 1e-9  # centred
    df, kf = fractal_fit(c.pos, a)
    assert abs(df - 1.8) < 0.05 and abs(kf - 1.3) < 0.15, (df, kf)
    # radius of gyration of the whole aggregate follows N = kf (Rg/a)^Df
    rg = math.sqrt((c.pos ** 2).sum(1).mean())
    assert n == pytest.approx(1.3 * (rg / a) ** 1.8, rel=0.1)


def test_fibonacci_directions():
    k = 32
    d = G.fibonacci_directions(k)
    cb = np.cos(d[:, 0])
    assert np.allclose(cb, 1 - 2 * (np.arange(k) + 0.5) / k)
    assert np.all((d[:, 1] >= 0) & (d[:, 1] < 2 * math.pi))
    golden = (1 + 5 ** 0.5) / 2
    assert np.allclose(d[:, 1], np.mod(2 * math.pi * np.arange(k) / golden, 2 * math.pi))
    u = np.stack([np.sin(d[:, 0]) * np.cos(d[:, 1]), np.sin(d[:, 0]) * np.sin(d[:, 1]), cb], 1)
    assert np.linalg.norm(u.sum(0)) / k < 0.05  # evenly spread: the mean direction nearly vanishes


# ------------------------------------------------------------ position files
def test_read_pos_column_counts(tmp_path):
    p = tmp_path / "s.pos"
    p.write_text(" 1.0 2.0 3.0\n"                          # 3 numbers: radius 1, index 1
                 "-1.5e+00 0.5 2.0 0.75\n"                 # 4: radius
                 "0 0 1 2.0 1.5 0.01\n"                    # 6: index
                 "0 0 2 2.0 1.5\n"                          # 5: the index needs both parts
                 "0 0 3 2.0 (1.4,0.02)\n"                  # list-directed complex
                 "0 0 4 2.0 1.5 0.01 1.6 0.03\n"           # 8: second (chiral) index
                 "0 0 5 2.0 1.5 0.01 1.6\n")               # 7: second index = first
    c, m2 = G.read_pos(str(p), with_chiral=True)
    assert c.n == 7
    np.testing.assert_array_equal(c.pos[1], [-1.5, 0.5, 2.0])
    np.testing.assert_array_equal(c.radius, [1.0, 0.75, 2.0, 2.0, 2.0, 2.0, 2.0])
    np.testing.assert_array_equal(c.ri, [1, 1, 1.5 + 0.01j, 1, 1.4 + 0.02j, 1.5 + 0.01j, 1.5 + 0.01j])
    np.testing.assert_array_equal(m2, [1, 1, 1.5 + 0.01j, 1, 1.4 + 0.02j, 1.6 + 0.03j, 1.5 + 0.01j])


def test_read_pos_scale_factors_and_count(tmp_path):
    p = tmp_path / "s.pos"
    p.write_text("1 2 3 0.5 1.5 0.01\n4 5 6 0.25\n7 8 9\n")
    c = G.read_pos(str(p), number_spheres=2, length_scale_factor=2.0, ref_index_scale_factor=1.0 / 1.33)
    assert c.n == 2  # reading stops at number_spheres
    np.testing.assert_allclose(c.pos, [[2, 4, 6], [8, 10, 12]])
    np.testing.assert_allclose(c.radius, [1.0, 0.5])
    np.testing.assert_allclose(c.ri, [(1.5 + 0.01j) / 1.33, 1 / 1.33])
    with pytest.raises(ValueError, match="insufficient data"):
        G.read_pos(str(p), number_spheres=4)
    bad = tmp_path / "b.pos"
    bad.write_text("1 2 3\n1 2\n")
    with pytest.raises(ValueError, match="read error"):
        G.read_pos(str(bad), number_spheres=2)


def test_write_read_round_trip(tmp_path):
    c = G.rsa_in_sphere(50, 1.3, 0.1, seed=5, ri=1.5 + 0.01j)
    p = tmp_path / "c.pos"
    G.write_pos(str(p), c)
    r = G.read_pos(str(p))
    np.testing.assert_array_equal(r.pos, c.pos)
    np.testing.assert_array_equal(r.radius, c.radius)
    np.testing.assert_array_equal(r.ri, c.ri)


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples not present")
def test_read_reference_example_files():
    # the two position files the reference ships: 4 numbers per line (unit
    # spheres) and 3 per line (radius 1 by default)
    cyl = G.read_pos(os.path.join(REF_EXAMPLES, "random_in_cylinder_1000.pos"))
    agg = G.read_pos(os.path.join(REF_EXAMPLES, "frac_agg.pos"))
    assert cyl.n == 1000 and np.all(cyl.radius == 1.0) and np.all(cyl.ri == 1.0)
    assert agg.n == 100 and np.all(agg.radius == 1.0)
    assert pair_min_gap(cyl.pos, 1.0) > -1e-4
