"""Host-side mirror of MSTM's solver interface over the C-ABI.

Method names follow the reference procedures they stand in for (solver and
fft_translation modules of MSTM's december2023/ tree); arrays use the same
layouts (coefficient vectors (nrhs, number_eqns) complex128 = Fortran
(number_eqns, nrhs)).  Everything runs in libmstm_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import FP32, FP64, SOLVE_FGMRES, SOLVE_LEARNED, SOLVE_MIXED, SOLVE_PLAIN, MstmHipError, load_library  # noqa: F401


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class HipContext:
    """One GPU context holding one cluster (spheredata/mie module state)."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.mstm_hip_create(C.byref(h), device)
        if rc != 0:
            raise MstmHipError(f"mstm_hip_create failed ({rc}): no usable GPU {device}")
        self.h = h
        self.nsphere = 0

    # -- plumbing -----------------------------------------------------------
    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self.L.mstm_hip_last_error(self.h)
            raise MstmHipError(f"{what}: code {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mstm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- spheredata / mie (miecoefcalc, mstm-modules-33.f90:5301-5407) -------
    def set_spheres(self, pos, radius, ri, ri_medium: float = 1.0, mie_epsilon: float = 1e-6):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        radius = np.ascontiguousarray(radius, np.float64)
        ri = np.ascontiguousarray(np.asarray(ri, np.complex128)).view(np.float64)
        self.nsphere = len(radius)
        self._chk(self.L.mstm_hip_set_spheres(self.h, self.nsphere, _d(pos), _d(radius), _d(ri), float(ri_medium),
                                              float(mie_epsilon)), "set_spheres")
        self.radius = radius.copy()
        return self.orders()

    def set_mie(self, order, an_mie):
        order = np.ascontiguousarray(order, np.int32)
        an = np.ascontiguousarray(an_mie, np.complex128)
        self._chk(self.L.mstm_hip_set_mie(self.h, _i(order), _d(an.view(np.float64))), "set_mie")

    def orders(self):
        o = np.zeros(self.nsphere, np.int32)
        off = np.zeros(self.nsphere + 1, np.int32)
        self._chk(self.L.mstm_hip_get_orders(self.h, _i(o), _i(off)), "get_orders")
        self.sphere_order, self.sphere_offset = o, off
        return o, off

    @property
    def number_eqns(self) -> int:
        return self.L.mstm_hip_number_eqns(self.h)

    def an_mie(self):
        n = 4 * int(self.sphere_order.sum())
        out = np.zeros(n, np.complex128)
        self._chk(self.L.mstm_hip_get_mie(self.h, _d(out.view(np.float64))), "get_mie")
        return out

    # -- operators -----------------------------------------------------------
    def setup_direct(self, precision: int = FP64):
        self._chk(self.L.mstm_hip_setup_direct(self.h, precision), "setup_direct")

    def node_selection(self, cell_volume_fraction: float = 0.0, node_order: int = -1, neighbor_node_model: int = 2):
        cd = np.zeros(3, np.int32)
        d = np.zeros(1)
        no = np.zeros(1, np.int32)
        self._chk(self.L.mstm_hip_node_selection(self.h, cell_volume_fraction, node_order, neighbor_node_model, _i(cd),
                                                 _d(d), _i(no)), "node_selection")
        return dict(cell_dim=cd, d_cell=float(d[0]), node_order=int(no[0]))

    def setup_fft(self, precision: int = FP64):
        self._chk(self.L.mstm_hip_setup_fft(self.h, precision), "setup_fft")

    def sphereinteraction(self, ain, con_tran, mie_mult=None, fft_option=False, precision: int = FP64):
        """aout = A ain per column (sphereinteraction, mstm-scatprops-26.f90:25-180)."""
        ain = np.ascontiguousarray(np.atleast_2d(ain), np.complex128)
        nrhs = ain.shape[0]
        ct = np.array([1 if x else 0 for x in con_tran], np.int32)
        mm = np.array([1 if x else 0 for x in (mie_mult if mie_mult is not None else [True] * nrhs)], np.int32)
        out = np.zeros_like(ain)
        self._chk(self.L.mstm_hip_sphereinteraction(self.h, nrhs, _d(ain.view(np.float64)), _d(out.view(np.float64)),
                                                    _i(ct), _i(mm), int(bool(fft_option)), precision),
                  "sphereinteraction")
        return out

    # -- solver (cbicg / fixedorsoln, mstm-solver-8.f90) ----------------------
    def cbicg(self, niter, eps, pnp, anp, fft_option=False):
        pnp = np.ascontiguousarray(pnp, np.complex128)
        x = np.ascontiguousarray(anp, np.complex128).copy()
        it = np.zeros(1, np.int32)
        err = np.zeros(1)
        self._chk(self.L.mstm_hip_cbicg(self.h, niter, eps, _d(pnp.view(np.float64)), _d(x.view(np.float64)),
                                        int(bool(fft_option)), _i(it), _d(err)), "cbicg")
        return x, int(it[0]), float(err[0])

    def fixedorsoln(self, alpha=0.0, sinc=0.0, eps=1e-6, niter=10000, solver=SOLVE_PLAIN, fft_option=False,
                    direction=1):
        ne = self.number_eqns
        amnp = np.zeros((2, ne), np.complex128)
        qeff = np.zeros((self.nsphere, 3, 3))
        maxerr = np.zeros(1)
        maxiter = np.zeros(1, np.int32)
        istat = np.zeros(1, np.int32)
        self._chk(self.L.mstm_hip_fixedorsoln(self.h, alpha, sinc, direction, eps, niter, solver, int(bool(fft_option)),
                                              _d(amnp.view(np.float64)), _d(qeff), _d(maxerr), _i(maxiter),
                                              _i(istat)), "fixedorsoln")
        return dict(amnp=amnp, qeff=qeff, maxerr=float(maxerr[0]), iterations=int(maxiter[0]), istat=int(istat[0]))

    # -- learned initial guess (no reference counterpart) ---------------------
    def set_learned(self, params=None, hidden: int = 128, nlayers: int = 3, lnet: int = 4):
        """Load a network for SOLVE_LEARNED (None unloads it)."""
        if params is None:
            self._chk(self.L.mstm_hip_set_learned(self.h, None, 0, 0, 0, 0), "set_learned")
            return
        p = np.ascontiguousarray(params, np.float32)
        self._lig_params = p
        self._chk(self.L.mstm_hip_set_learned(self.h, p.ctypes.data_as(C.POINTER(C.c_float)), len(p), hidden, nlayers,
                                              lnet), "set_learned")

    def load_learned(self, path: str | None = None):
        """Load a weight file written by mstm_amd.learned (default: the shipped one)."""
        from . import learned

        p, hidden, nlayers, lnet, meta = learned.load_weights(path or learned.DEFAULT_WEIGHTS)
        self.set_learned(p, hidden, nlayers, lnet)
        return meta

    def lig_features_into(self, feat, depth):
        self._chk(self.L.mstm_hip_lig_features(self.h, feat.ctypes.data_as(C.POINTER(C.c_float)),
                                               depth.ctypes.data_as(C.POINTER(C.c_float))), "lig_features")

    def lig_features(self):
        feat = np.zeros((self.nsphere, 64), np.float32)
        depth = np.zeros(self.nsphere, np.float32)
        self.lig_features_into(feat, depth)
        return feat, depth

    def lig_samples_into(self, amnp, lnet, target, base):
        a = np.ascontiguousarray(amnp, np.complex128)
        self._chk(self.L.mstm_hip_lig_samples(self.h, _d(a.view(np.float64)), lnet,
                                              target.ctypes.data_as(C.POINTER(C.c_float)),
                                              base.ctypes.data_as(C.POINTER(C.c_float))), "lig_samples")

    def lig_guess(self):
        x0 = np.zeros((2, self.number_eqns), np.complex128)
        self._chk(self.L.mstm_hip_lig_guess(self.h, _d(x0.view(np.float64))), "lig_guess")
        return x0

    def set_inner_solver(self, kind: int):
        """Inner solver of the mixed refinement: 0 BiCG, 1 GMRES (default)."""
        self._chk(self.L.mstm_hip_set_inner_solver(self.h, int(kind)), "set_inner_solver")

    def clear(self, clear_tables=True):
        self._chk(self.L.mstm_hip_clear(self.h, int(bool(clear_tables))), "clear")

    def stats(self) -> dict:
        names = self.L.mstm_hip_stats_names().decode().split(",")
        out = np.zeros(len(names))
        k = self.L.mstm_hip_stats(self.h, _d(out), len(names))
        return dict(zip(names[:k], out[:k].tolist()))


def qtotcalc(radius, qeff):
    """qtotcalc (mstm-scatprops-26.f90:872-912) for external spheres, with the
    cross-section radius of main_calling_program (mstm-input-37.f90:1204-1211):
    (sum a^3) raised to the real(4) literal .333333.  Returns qtot[k, q],
    k = unpolarized/pol1/pol2, q = ext/abs/sca."""
    radius = np.asarray(radius, np.float64)
    xg = float(np.sum(radius ** 3)) ** float(np.float32(0.333333))
    w = radius ** 2 / xg ** 2
    q = np.einsum("i,ikq->kq", w, np.asarray(qeff))
    q[:, 2] = q[:, 0] + q[:, 2] - q[:, 1]
    return q
