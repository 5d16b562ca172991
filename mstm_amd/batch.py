"""Batched solve of many clusters x many incidence directions on one GPU
(libmstm_hip.so, mstm_hip_batch_*).

MSTM solves each cluster and each incidence separately (fixedorsoln,
mstm-solver-8.f90:350-569).  BatchSolver takes a list of clusters and one list
of incidence directions; every (cluster, direction, polarization) is solved to
the reference criterion |r|^2/|b|^2 < eps.  Host vectors are per cluster
(R, neqns_c) complex arrays in MSTM's coefficient layout, R = 2 x directions
(column 2d + q: direction d, polarization q).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import geometry as G
from .lib import FP32, FP64, MstmHipError, load_library  # noqa: F401


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def random_clusters(ncluster: int = 256, seed: int = 3, nmin: int = 10, nmax: int = 300, xmin: float = 0.5,
                    xmax: float = 3.0, fvmin: float = 0.05, fvmax: float = 0.30, ri: complex = 1.33 + 0.001j):
    """Clusters of equal spheres: N log-uniform in [nmin, nmax], size parameter
    uniform in [xmin, xmax], volume fraction uniform in [fvmin, fvmax], seeded
    random sequential addition inside a sphere (one seed per cluster)."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncluster))).astype(int)
    xs = rng.uniform(xmin, xmax, ncluster)
    fvs = rng.uniform(fvmin, fvmax, ncluster)
    return [G.rsa_in_sphere(int(n), float(x), float(fv), seed=seed * 100003 + k, ri=ri)
            for k, (n, x, fv) in enumerate(zip(ns, xs, fvs))]


class BatchSolver:
    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.mstm_hip_batch_create(C.byref(h), device)
        if rc != 0:
            raise MstmHipError(f"mstm_hip_batch_create failed ({rc}): no usable GPU {device}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.mstm_hip_batch_last_error(self.h)
            raise MstmHipError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mstm_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, clusters, directions, ri_medium: float = 1.0, mie_epsilon: float = 1e-6):
        """clusters: list of geometry.Cluster; directions: (ndir, 2) array of (beta, alpha) radians."""
        self.clusters = clusters
        counts = np.array([c.n for c in clusters], np.int32)
        pos = np.ascontiguousarray(np.concatenate([c.pos for c in clusters]), np.float64)
        rad = np.ascontiguousarray(np.concatenate([c.radius for c in clusters]), np.float64)
        ri = np.ascontiguousarray(np.concatenate([np.asarray(c.ri, np.complex128) for c in clusters])).view(np.float64)
        d = np.asarray(directions, np.float64).reshape(-1, 2)
        beta = np.ascontiguousarray(d[:, 0])
        alpha = np.ascontiguousarray(d[:, 1])
        self.ndir = len(d)
        self.R = 2 * self.ndir
        self._chk(self.L.mstm_hip_batch_set(self.h, len(clusters), _i(counts), _d(pos), _d(rad), _d(ri),
                                            float(ri_medium), float(mie_epsilon), self.ndir, _d(alpha), _d(beta)),
                  "batch_set")
        self.order = np.zeros(len(clusters), np.int32)
        self._chk(self.L.mstm_hip_batch_orders(self.h, _i(self.order)), "batch_orders")
        self.neqns = [int(2 * c.n * L * (L + 2)) for c, L in zip(clusters, self.order)]
        self.total = self.R * sum(self.neqns)

    def setup(self, tune: bool = True):
        self._chk(self.L.mstm_hip_batch_setup(self.h, int(bool(tune))), "batch_setup")
        st = self.stats()
        self.nrows = int(st["nrows"])

    def solve(self, eps: float = 1e-6, niter: int = 10000, use_graph: bool = True):
        n = len(self.clusters) * self.R
        it = np.zeros(n, np.int32)
        err = np.zeros(n)
        self._chk(self.L.mstm_hip_batch_solve(self.h, float(eps), int(niter), int(bool(use_graph)), _i(it), _d(err)),
                  "batch_solve")
        return dict(iterations=it.reshape(-1, self.R), err=err.reshape(-1, self.R))


    def _split(self, flat):
        out, o = [], 0
        for ne in self.neqns:
            out.append(flat[o:o + self.R * ne].reshape(self.R, ne))
            o += self.R * ne
        return out

    def vectors(self, which: int = 0):
        """0: solution, 1: right-hand side T f, 2: incident field f; list of (R, neqns_c)."""
        flat = np.zeros(self.total, np.complex128)
        self._chk(self.L.mstm_hip_batch_vectors(self.h, which, _d(flat.view(np.float64))), "batch_vectors")
        return self._split(flat)

    def product(self, xs, precision: int = FP64, op: int = 0):
        flat = np.ascontiguousarray(np.concatenate([np.asarray(x, np.complex128).ravel() for x in xs]))
        out = np.zeros_like(flat)
        self._chk(self.L.mstm_hip_batch_product(self.h, precision, op, _d(flat.view(np.float64)),
                                                _d(out.view(np.float64))), "batch_product")
        return self._split(out)

    def qeff(self):
        q = np.zeros((len(self.clusters), self.ndir, 3, 3))
        self._chk(self.L.mstm_hip_batch_qeff(self.h, _d(q)), "batch_qeff")
        return q

    def stats(self) -> dict:
        names = self.L.mstm_hip_batch_stats_names().decode().split(",")
        out = np.zeros(len(names))
        k = self.L.mstm_hip_batch_stats(self.h, _d(out), len(names))
        return dict(zip(names[:k], out[:k].tolist()))

    def tune_table(self):
        out = np.zeros(300)
        k = self.L.mstm_hip_batch_tune_table(self.h, _d(out), len(out))
        return out[:k].reshape(-1, 3)
