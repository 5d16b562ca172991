"""Learned initial guess: training data from the solver itself, training on
the GPU through the library's trainer (hand-written HIP forward, backward and
Adam kernels), and the weight files the learned solver loads.

The model, shared with libmstm_hip.so: for incidence along +z the solution
block of sphere i starts at

    x0_i = exp(i kappa d_i) * (b_i + exp(i k z_i) * NN(f_i))

b_i = T_i f_i is MSTM's default start (mstm-solver-8.f90:483), d_i the depth
below the cluster surface along the incidence direction, kappa a learned
complex wavenumber shift and NN a per-sphere MLP (64 neighbour-geometry
features, 3 SiLU layers of 128, linear output of both polarizations of an
order-4 sphere).  Training pairs are (features, a_i exp(-i k z_i)) from
converged solves of random clusters.

Command line (runs on the GPU):
    python -m mstm_amd.learned train --configs 2000 --out mstm_amd/weights/lig_x2_m131_fv030.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import geometry as G
from .lib import SOLVE_MIXED, MstmHipError, load_library

HIDDEN, NLAYERS, LNET = 128, 3, 4
NFEAT = 64
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "lig_x2_m131_fv030.npz")


def nout(lnet: int = LNET) -> int:
    return 8 * lnet * (lnet + 2)


def nparams(hidden: int = HIDDEN, nlayers: int = NLAYERS, lnet: int = LNET) -> int:
    return int(load_library().mstm_hip_lig_nparams(hidden, nlayers, lnet))


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- weights
def save_weights(path: str, params, hidden: int = HIDDEN, nlayers: int = NLAYERS, lnet: int = LNET, meta=None):
    """path (.npz: params, shape, meta) and path with .bin: the raw form a
    Fortran host reads (int32 hidden, nlayers, lnet, nparams; float32 params)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    p = np.asarray(params, np.float32)
    np.savez(path, params=p, shape=np.array([hidden, nlayers, lnet], np.int32), meta=np.array(json.dumps(meta or {})))
    with open(os.path.splitext(path)[0] + ".bin", "wb") as fh:
        fh.write(np.array([hidden, nlayers, lnet, len(p)], np.int32).tobytes())
        fh.write(p.tobytes())


def load_weights(path: str = DEFAULT_WEIGHTS):
    """(params float32, hidden, nlayers, lnet, meta dict); plain arrays only."""
    with np.load(path, allow_pickle=False) as z:
        hidden, nlayers, lnet = (int(v) for v in z["shape"])
        return z["params"].astype(np.float32), hidden, nlayers, lnet, json.loads(str(z["meta"]))


# ---------------------------------------------------------------- trainer
class Trainer:
    """The library's MLP trainer (mstm_hip_trainer_*)."""

    def __init__(self, device: int = 0, hidden: int = HIDDEN, nlayers: int = NLAYERS, lnet: int = LNET,
                 seed: int = 1):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.mstm_hip_trainer_create(C.byref(h), device, hidden, nlayers, lnet, seed)
        self.h = h
        self._chk(rc, "trainer_create")
        self.hidden, self.nlayers, self.lnet = hidden, nlayers, lnet
        self.np = nparams(hidden, nlayers, lnet)

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.mstm_hip_trainer_last_error(self.h) if self.h else b""
            raise MstmHipError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mstm_hip_trainer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.np, np.float32)
        self._chk(self.L.mstm_hip_trainer_params(self.h, _f(p), self.np, 0), "params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(self.L.mstm_hip_trainer_params(self.h, _f(p), len(p), 1), "params")

    def set_data(self, feat, target, base, depth):
        feat, target, base, depth = (np.ascontiguousarray(a, np.float32) for a in (feat, target, base, depth))
        self._keep = (feat, target, base, depth)
        self._chk(self.L.mstm_hip_trainer_set_data(self.h, len(depth), _f(feat), _f(target), _f(base), _f(depth)),
                  "set_data")
        self._keep = None

    def epoch(self, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        l = np.zeros(1)
        self._chk(self.L.mstm_hip_trainer_epoch(self.h, batch, C.c_float(lr), seed, _d(l)), "epoch")
        return float(l[0])

    def loss(self, batch: int = 8192) -> float:
        l = np.zeros(1)
        self._chk(self.L.mstm_hip_trainer_loss(self.h, batch, _d(l)), "loss")
        return float(l[0])

    def grad(self, feat, target, base, depth):
        feat, target, base, depth = (np.ascontiguousarray(a, np.float32) for a in (feat, target, base, depth))
        g = np.zeros(self.np, np.float32)
        l = np.zeros(1)
        self._chk(self.L.mstm_hip_trainer_grad(self.h, len(depth), _f(feat), _f(target), _f(base), _f(depth), _f(g),
                                               _d(l)), "grad")
        return g, float(l[0])

    def forward(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((len(feat), nout(self.lnet)), np.float32)
        self._chk(self.L.mstm_hip_trainer_forward(self.h, len(feat), _f(feat), _f(out)), "forward")
        return out


# ---------------------------------------------------------------- data
def _make_cluster(args):
    n, seed, a, ri, fv = args
    return G.rsa_in_sphere(n, a, fv, seed=seed, ri=ri)


def cluster_jobs(nconfig: int, nmin: int, nmax: int, seed: int, a: float = 2.0, ri: complex = 1.31,
                 fv: float = 0.30):
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, nconfig)
    return [(int(n), int(seed * 100000 + k), a, ri, fv) for k, n in enumerate(ns)]


def cluster_samples(ctx, amnp, lnet: int = LNET):
    """(feat, depth, target, base) of every sphere of the context's cluster."""
    n = ctx.nsphere
    feat = np.zeros((n, NFEAT), np.float32)
    depth = np.zeros(n, np.float32)
    tgt = np.zeros((n, nout(lnet)), np.float32)
    base = np.zeros((n, nout(lnet)), np.float32)
    ctx.lig_features_into(feat, depth)
    ctx.lig_samples_into(amnp, lnet, tgt, base)
    return feat, depth, tgt, base


def generate(jobs, eps: float = 1e-6, workers: int = 12, log=print):
    """Solve every cluster of `jobs` with the mixed solver; returns stacked samples."""
    from .solver import HipContext

    F, D, T, B = [], [], [], []
    t0 = time.time()
    with ProcessPoolExecutor(workers) as ex:
        # every job is submitted (and every worker started) before this
        # process touches the GPU
        clusters = ex.map(_make_cluster, jobs, chunksize=2)
        ctx = HipContext(0)
        for k, c in enumerate(clusters):
            ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
            r = ctx.fixedorsoln(eps=eps, niter=10000, solver=SOLVE_MIXED)
            if r["istat"] != 0:
                log(f"  config {k}: not converged, skipped")
                continue
            f, d, t, b = cluster_samples(ctx, r["amnp"])
            keep = ctx.sphere_order == LNET
            F.append(f[keep]), D.append(d[keep]), T.append(t[keep]), B.append(b[keep])
            if k % 50 == 0 or k == len(jobs) - 1:
                log(f"  config {k + 1}/{len(jobs)}: {len(c.radius)} spheres, {r['iterations']} iterations, "
                    f"{time.time() - t0:.0f} s")
    ctx.close()
    return np.concatenate(F), np.concatenate(D), np.concatenate(T), np.concatenate(B)


def train(feat, depth, target, base, epochs: int = 40, batch: int = 4096, lr: float = 2e-3, seed: int = 1,
          log=print, trainer=None):
    tr = trainer or Trainer(0, seed=seed)
    tr.set_data(feat, target, base, depth)
    l0 = tr.loss()
    log(f"  initial loss {l0:.4f}")
    hist = [l0]
    for e in range(epochs):
        lr_e = lr * 0.5 * (1.0 + np.cos(np.pi * e / epochs))  # cosine schedule
        l = tr.epoch(batch, lr_e, seed=seed * 1000 + e)
        hist.append(l)
        if e % 5 == 0 or e == epochs - 1:
            log(f"  epoch {e + 1}/{epochs}: loss {l:.4f} (lr {lr_e:.2e})")
    return tr, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--configs", type=int, default=2000)
    t.add_argument("--nmin", type=int, default=200)
    t.add_argument("--nmax", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    t0 = time.time()
    log(f"generating {a.configs} clusters of {a.nmin}-{a.nmax} spheres (x=2, m=1.31, fv=0.30)")
    f, d, tg, b = generate(cluster_jobs(a.configs, a.nmin, a.nmax, a.seed), log=log)
    log(f"{len(d)} samples in {time.time() - t0:.0f} s; training")
    tr, hist = train(f, d, tg, b, epochs=a.epochs, seed=a.seed, log=log)
    p = tr.params
    meta = dict(configs=a.configs, nmin=a.nmin, nmax=a.nmax, seed=a.seed, epochs=a.epochs, samples=int(len(d)),
                loss_initial=hist[0], loss_final=hist[-1], kappa=[float(p[-2]), float(p[-1])],
                size_parameter=2.0, ref_index=[1.31, 0.0], volume_fraction=0.30)
    save_weights(a.out, p, meta=meta)
    log(f"saved {a.out}: kappa {p[-2]:.5f}{p[-1]:+.5f}i, loss {hist[0]:.4f} -> {hist[-1]:.4f}, "
        f"{time.time() - t0:.0f} s")


if __name__ == "__main__":
    sys.exit(main())
