"""Seeded cluster geometries and the MSTM position-file format.

The reference reads sphere data as lines ``x y z [r [m [m2]]]``
(read_sphere_data_input_file, mstm-input-37.f90:2871-2924), each complex
index either as two numbers or as ``(re,im)``; missing radius -> 1, missing
index -> 1+0i, missing second (chiral) index -> the first.  Positions and
radii are then scaled by length_scale_factor and both indices by
ref_index_scale_factor.  Its own random generator is unseeded
(random_configuration-10.f90:77-80), so the workloads here come from these
seeded samplers instead.  Lengths are in units of 1/k (size parameter units),
as with length_scale_factor = 1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class Cluster:
    pos: np.ndarray  # (N, 3) float64
    radius: np.ndarray  # (N,) float64
    ri: np.ndarray  # (N,) complex128
    ri_medium: complex = 1.0 + 0.0j

    @property
    def n(self) -> int:
        return len(self.radius)


def write_pos(path: str, c: Cluster) -> None:
    with open(path, "w") as f:
        for p, r, m in zip(c.pos, c.radius, c.ri):
            f.write(f"{p[0]:.17e} {p[1]:.17e} {p[2]:.17e} {r:.17e} {m.real:.17e} {m.imag:.17e}\n")


def _pos_fields(line: str) -> list[float]:
    """The numbers of one position-file line, a parenthesised complex
    ``(re,im)`` counting as its two parts, as list-directed input reads it."""
    out = []
    for tok in line.replace("(", " ").replace(")", " ").replace(",", " ").split():
        out.append(float(tok.replace("d", "e").replace("D", "e")))
    return out


def read_pos(path: str, number_spheres: int | None = None, length_scale_factor: float = 1.0,
             ref_index_scale_factor: complex = 1.0, with_chiral: bool = False):
    """read_sphere_data_input_file (mstm-input-37.f90:2871-2924): up to
    number_spheres lines (all when None) of ``x y z [r [m_re m_im [m2_re
    m2_im]]]``.  A line gives the radius from 4 numbers on, the index from 6
    on and the second index from 8 on; otherwise radius 1, index 1+0i and the
    second index equal to the first.  A line with fewer than three numbers,
    or fewer lines than number_spheres, is an error, as in the reference.
    Returns a Cluster (and the (N,) second indices when with_chiral)."""
    pos, rad, ri, ri2 = [], [], [], []
    with open(path) as f:
        for line in f:
            if number_spheres is not None and len(pos) == number_spheres:
                break
            v = _pos_fields(line)
            if number_spheres is None and not v:
                continue
            if len(v) < 3:
                raise ValueError(f"read error in sphere data input file {path}, line {len(pos) + 1}")
            pos.append(v[:3])
            rad.append(v[3] if len(v) >= 4 else 1.0)
            m1 = complex(v[4], v[5]) if len(v) >= 6 else 1.0 + 0.0j
            ri.append(m1)
            ri2.append(complex(v[6], v[7]) if len(v) >= 8 else m1)
    if number_spheres is not None and len(pos) < number_spheres:
        raise ValueError(f"insufficient data in input file: {len(pos)} lines, need {number_spheres}")
    ls, rs = float(length_scale_factor), complex(ref_index_scale_factor)
    c = Cluster(np.array(pos, dtype=float) * ls, np.array(rad, dtype=float) * ls, np.array(ri, dtype=complex) * rs)
    return (c, np.array(ri2, dtype=complex) * rs) if with_chiral else c


class _Grid:
    """Uniform hash grid for overlap tests of equal spheres."""

    def __init__(self, cell: float):
        self.cell = cell
        self.cells: dict[tuple[int, int, int], list[int]] = {}

    def key(self, p):
        return (int(math.floor(p[0] / self.cell)), int(math.floor(p[1] / self.cell)), int(math.floor(p[2] / self.cell)))

    def add(self, idx: int, p) -> None:
        self.cells.setdefault(self.key(p), []).append(idx)

    def near(self, p):
        kx, ky, kz = self.key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    yield from self.cells.get((kx + dx, ky + dy, kz + dz), ())


def _rsa(n: int, a: float, sample, rng: np.random.Generator, max_tries: int = 2_000_000) -> np.ndarray:
    """Random sequential addition of n spheres of radius a (touching allowed)."""
    pos = np.zeros((n, 3))
    grid = _Grid(2.0 * a)
    d2 = (2.0 * a) ** 2
    k = 0
    tries = 0
    batch = 4096
    while k < n:
        cand = sample(rng, batch)
        for p in cand:
            tries += 1
            ok = True
            for j in grid.near(p):
                q = pos[j]
                if (p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2 + (p[2] - q[2]) ** 2 < d2:
                    ok = False
                    break
            if ok:
                pos[k] = p
                grid.add(k, p)
                k += 1
                if k == n:
                    break
        if tries > max_tries:
            raise RuntimeError(f"RSA jammed after {k} of {n} spheres")
    return pos


def rsa_in_sphere(n: int, a: float, fv: float, seed: int, ri: complex, ri_medium: complex = 1.0) -> Cluster:
    """n spheres of radius a, centres uniform in |r| <= R - a, R = a (n/fv)^(1/3).
    Small clusters at high fv can jam (few centres fit in R - a): the bounding
    radius then grows by 3% per retry, same seed, until the addition completes."""
    R = a * (n / fv) ** (1.0 / 3.0)
    for attempt in range(40):
        rin = R - a

        def sample(rng, m):
            v = rng.normal(size=(m, 3))
            v /= np.linalg.norm(v, axis=1)[:, None]
            return v * (rin * rng.random(m) ** (1.0 / 3.0))[:, None]

        try:
            pos = _rsa(n, a, sample, np.random.default_rng(seed), max_tries=2_000_000 if attempt == 0 else 200_000)
            break
        except RuntimeError:
            R *= 1.03
    else:
        raise RuntimeError(f"RSA jammed for n={n}, fv={fv}")
    return Cluster(pos, np.full(n, a), np.full(n, ri, dtype=complex), ri_medium)


def rsa_in_cylinder(n: int, a: float, fv: float, seed: int, ri: complex, ri_medium: complex = 1.0) -> Cluster:
    """Cylinder along z of radius Rc and length 4 Rc (= 2 diameters), fv = n v_a / V."""
    Rc = a * (n / (3.0 * fv)) ** (1.0 / 3.0)
    Lz = 4.0 * Rc

    def sample(rng, m):
        rr = (Rc - a) * np.sqrt(rng.random(m))
        th = 2.0 * np.pi * rng.random(m)
        z = (Lz - 2.0 * a) * (rng.random(m) - 0.5)
        return np.stack([rr * np.cos(th), rr * np.sin(th), z], axis=1)

    pos = _rsa(n, a, sample, np.random.default_rng(seed))
    return Cluster(pos, np.full(n, a), np.full(n, ri, dtype=complex), ri_medium)


def fractal_aggregate(n: int, a: float, df: float, kf: float, seed: int, ri: complex, ri_medium: complex = 1.0,
                      max_tries: int = 20000) -> Cluster:
    """Sequential particle-cluster aggregation of touching spheres that keeps
    N = kf (Rg/a)^df at every step (radius of gyration of the growing cluster)."""
    rng = np.random.default_rng(seed)
    pos = [np.zeros(3)]
    grid = _Grid(2.0 * a)
    grid.add(0, pos[0])
    d2 = (2.0 * a) ** 2 * (1.0 - 1e-12)
    for k in range(1, n):
        P = np.array(pos)
        cm = P.mean(axis=0)
        rg2 = ((P - cm) ** 2).sum(axis=1).mean() if k > 1 else 0.0
        # target distance of the new monomer from the centre of mass
        N1 = k + 1
        rg_new2 = a * a * (N1 / kf) ** (2.0 / df)
        gam2 = (N1 * N1 * rg_new2 - N1 * k * rg2 - N1 * 0.0) / k if k > 1 else (2 * a) ** 2
        gam2 = max(gam2, 0.0)
        placed = False
        for _ in range(max_tries):
            # pick a member and a direction so the new monomer touches it
            j = int(rng.integers(k))
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            p = P[j] + 2.0 * a * v
            dist2 = ((p - cm) ** 2).sum()
            if k > 1 and abs(dist2 - gam2) > 0.25 * gam2 + (2 * a) ** 2:
                continue
            ok = True
            for q in grid.near(p):
                if q != j and ((p - P[q]) ** 2).sum() < d2:
                    ok = False
                    break
            if ok:
                pos.append(p)
                grid.add(k, p)
                placed = True
                break
        if not placed:
            raise RuntimeError(f"aggregation stalled at monomer {k}")
    P = np.array(pos)
    P -= P.mean(axis=0)
    return Cluster(P, np.full(n, a), np.full(n, ri, dtype=complex), ri_medium)


def fibonacci_directions(k: int) -> np.ndarray:
    """(beta, alpha) of k Fibonacci-sphere directions: cos(beta_i) = 1 - 2(i+1/2)/k,
    alpha_i = 2 pi i / golden ratio."""
    i = np.arange(k)
    cb = 1.0 - 2.0 * (i + 0.5) / k
    golden = (1.0 + 5.0 ** 0.5) / 2.0
    alpha = np.mod(2.0 * np.pi * i / golden, 2.0 * np.pi)
    return np.stack([np.arccos(cb), alpha], axis=1)
