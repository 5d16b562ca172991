"""Host-side mirror of MSTM's solver interface over the C-ABI.

Method names follow the reference procedures they stand in for (solver and
fft_translation modules of MSTM's december2023/ tree); arrays use the same
layouts (coefficient vectors (nrhs, number_eqns) complex128 = Fortran
(number_eqns, nrhs)).  Everything runs in libmstm_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import FP32, FP64, FP64_FP32, SOLVE_FGMRES, SOLVE_LEARNED, SOLVE_LU, SOLVE_MIXED, SOLVE_PLAIN, MstmHipError, load_library  # noqa: F401


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# the reference's pi is a real(4) literal widened to real(8) (mstm-modules-33.f90:18)
PI_R4 = float(np.float32(np.pi))


def scattering_angles(alpha: float = 0.0):
    """Default scattering-matrix directions of a single-origin run without
    azimuthal averaging (mstm-input-37.f90:1339-1351 and 3036-3046): theta =
    -180..180 deg in 1 deg steps, cos theta = +-1 moved to +-0.9999999, and
    negative theta in the half-plane phi = alpha + pi.  Returns (theta_deg,
    cos theta, phi)."""
    th = np.arange(-180, 181)
    ct = np.cos(PI_R4 / 180.0 * th.astype(np.float64))
    ct[ct == 1.0] = 0.9999999
    ct[ct == -1.0] = -0.9999999
    return th, ct, np.where(th < 0, alpha + PI_R4, alpha)


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

    def node_selection(self, cell_volume_fraction: float = 0.0, node_order: int = -1, neighbor_node_model: int = 2,
                       target_min=None, target_max=None, d_specified: bool = False, d_cell: float = 0.0,
                       local_host: int = 0):
        """node_selection(fva, target_min, target_max, d_specified, local_host)
        (fft_translation-5.f90:654-825) and the node_order rule
        (mstm-input-37.f90:1256-1260).  Returns the grid: cell_dim, d_cell,
        node_order, the volume fraction used (fva), cell_boundary,
        sphere_node (nsphere, 3; 1-based) and number_neighbor_nodes."""
        cd = np.zeros(3, np.int32)
        d = np.array([float(d_cell)])
        fv = np.array([float(cell_volume_fraction)])
        no = np.zeros(1, np.int32)
        tmn = None if target_min is None else np.ascontiguousarray(target_min, np.float64)
        tmx = None if target_max is None else np.ascontiguousarray(target_max, np.float64)
        self._chk(self.L.mstm_hip_node_selection(self.h, _d(fv), None if tmn is None else _d(tmn),
                                                 None if tmx is None else _d(tmx), int(bool(d_specified)), _d(d),
                                                 int(local_host), int(node_order), int(neighbor_node_model), _i(cd),
                                                 _i(no)), "node_selection")
        cb = np.zeros(3)
        sn = np.zeros((self.nsphere, 3), np.int32)
        nn = np.zeros(1, np.int32)
        self._chk(self.L.mstm_hip_fft_grid(self.h, _d(cb), _i(sn), _i(nn)), "fft_grid")
        return dict(cell_dim=cd, d_cell=float(d[0]), node_order=int(no[0]), fva=float(fv[0]), cell_boundary=cb,
                    sphere_node=sn, number_neighbor_nodes=int(nn[0]))

    def setup_fft(self, precision: int = FP64):
        self._chk(self.L.mstm_hip_setup_fft(self.h, precision), "setup_fft")

    def sphereinteraction(self, ain, con_tran, mie_mult=None, fft_option=False, precision: int = FP64,
                          rhs_list=None, skip_external_translation: bool = False):
        """aout = A ain per column (sphereinteraction, mstm-scatprops-26.f90:25-180).
        Columns left out of rhs_list, or all with skip_external_translation,
        come back zero (no translation)."""
        ain = np.ascontiguousarray(np.atleast_2d(ain), np.complex128)
        nrhs = ain.shape[0]
        ct = np.array([1 if x else 0 for x in con_tran], np.int32)
        mm = np.array([1 if x else 0 for x in (mie_mult if mie_mult is not None else [True] * nrhs)], np.int32)
        rl = None if rhs_list is None else np.array([1 if x else 0 for x in rhs_list], np.int32)
        out = np.zeros_like(ain)
        self._chk(self.L.mstm_hip_sphereinteraction(self.h, nrhs, _d(ain.view(np.float64)), _d(out.view(np.float64)),
                                                    _i(ct), _i(mm), int(bool(fft_option)), precision,
                                                    None if rl is None else _i(rl),
                                                    int(bool(skip_external_translation))), "sphereinteraction")
        return out

    def sphereinteraction_noalloc(self, ain, con_tran, mie_mult=None, precision: int = FP64, rhs_list=None,
                                  skip_external_translation: bool = False):
        """The direct-operator product of sphereinteraction without device
        allocation (the library's fixed per-GPU scratch); the tables of this
        precision must exist (setup_direct)."""
        ain = np.ascontiguousarray(np.atleast_2d(ain), np.complex128)
        nrhs = ain.shape[0]
        ct = np.array([1 if x else 0 for x in con_tran], np.int32)
        mm = np.array([1 if x else 0 for x in (mie_mult if mie_mult is not None else [True] * nrhs)], np.int32)
        rl = None if rhs_list is None else np.array([1 if x else 0 for x in rhs_list], np.int32)
        out = np.zeros_like(ain)
        self._chk(self.L.mstm_hip_sphereinteraction_noalloc(self.h, nrhs, _d(ain.view(np.float64)),
                                                            _d(out.view(np.float64)), _i(ct), _i(mm), 0, precision,
                                                            None if rl is None else _i(rl),
                                                            int(bool(skip_external_translation))),
                  "sphereinteraction_noalloc")
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
                    direction=1, excited_spheres=None):
        """fixedorsoln (mstm-solver-8.f90:350-569).  niter < 0 runs the
        order-of-scattering series, niter = 0 no solve; excited_spheres
        (nsphere flags, None = all) switches the incident field per sphere."""
        ne = self.number_eqns
        amnp = np.zeros((2, ne), np.complex128)
        qeff = np.zeros((self.nsphere, 3, 3))
        maxerr = np.zeros(1)
        maxiter = np.zeros(1, np.int32)
        istat = np.zeros(1, np.int32)
        ex = None if excited_spheres is None else np.array([1 if v else 0 for v in excited_spheres], np.int32)
        if ex is not None and len(ex) != self.nsphere:
            raise ValueError("excited_spheres needs one flag per sphere")
        self._chk(self.L.mstm_hip_fixedorsoln(self.h, alpha, sinc, direction, eps, niter, solver, int(bool(fft_option)),
                                              _d(amnp.view(np.float64)), _d(qeff), _d(maxerr), _i(maxiter),
                                              _i(istat), None if ex is None else _i(ex)), "fixedorsoln")
        return dict(amnp=amnp, qeff=qeff, maxerr=float(maxerr[0]), iterations=int(maxiter[0]), istat=int(istat[0]))

    # -- post-solve observables of a single-origin expansion ------------------
    def translation_orders(self, origin=None, eps: float = 1e-5, maxt: int = 100):
        """(T, torder): translation orders about origin (default: the mean
        sphere position) as main_calling_program sets them
        (mstm-input-37.f90:1159-1197, tranordertest)."""
        t = np.zeros(self.nsphere, np.int32)
        T = np.zeros(1, np.int32)
        o = None if origin is None else np.ascontiguousarray(origin, np.float64)
        self._chk(self.L.mstm_hip_translation_orders(self.h, None if o is None else _d(o), float(eps), int(maxt),
                                                     _i(t), _i(T)), "translation_orders")
        self.t_matrix_order = int(T[0])
        return self.t_matrix_order, t

    def set_translation_orders(self, origin, torder, T: int):
        """Use a host's own translation_order / t_matrix_order / cluster_origin."""
        o = np.ascontiguousarray(origin, np.float64)
        t = np.ascontiguousarray(torder, np.int32)
        self._chk(self.L.mstm_hip_set_translation_orders(self.h, _d(o), _i(t), int(T)), "set_translation_orders")
        self.t_matrix_order = int(T)

    def merge_to_common_origin(self, amnp=None, copy: bool = True):
        """merge_to_common_origin (mstm-scatprops-26.f90:447-559) of both
        columns of a solution (None: the last fixedorsoln solution, still on
        the device).  Returns the E/M common-origin coefficients as a
        (2 columns, 2, T(T+2)) complex array, or None with copy=False (they
        stay on the device for scattering_matrix / asymmetry)."""
        T = self.t_matrix_order
        a = None if amnp is None else np.ascontiguousarray(amnp, np.complex128)
        out = np.zeros((2, 2, T * (T + 2)), np.complex128) if copy else None
        self._chk(self.L.mstm_hip_merge_to_common_origin(
            self.h, None if a is None else _d(a.view(np.float64)),
            None if out is None else _d(out.view(np.float64))), "merge_to_common_origin")
        return out

    def scattering_matrix(self, ct, phi, a0=None, T: int = 0):
        """scatteringmatrix (mstm-scatprops-26.f90:1498-1573) at each
        (cos theta, phi): (nang, 16) Mueller elements, Fortran sm(4,4)
        column-major per direction (S11 = [:, 0], S12 = [:, 4])."""
        ct = np.ascontiguousarray(np.atleast_1d(ct), np.float64)
        phi = np.ascontiguousarray(np.broadcast_to(np.atleast_1d(phi), ct.shape), np.float64)
        out = np.zeros((len(ct), 16))
        a = None if a0 is None else np.ascontiguousarray(a0, np.complex128)
        self._chk(self.L.mstm_hip_scattering_matrix(self.h, int(T), None if a is None else _d(a.view(np.float64)),
                                                    len(ct), _d(ct), _d(phi), _d(out)), "scattering_matrix")
        return out

    def asymmetry(self, a0=None, T: int = 0):
        """(s00(1,1,0:1), g): first two Legendre moments of S11 of
        fosmexpansion and g = s00(1,1,1) / (3 s00(1,1,0))."""
        s = np.zeros(2)
        g = np.zeros(1)
        a = None if a0 is None else np.ascontiguousarray(a0, np.complex128)
        self._chk(self.L.mstm_hip_asymmetry(self.h, int(T), None if a is None else _d(a.view(np.float64)), _d(s),
                                            _d(g)), "asymmetry")
        return s, float(g[0])

    def fosmexpansion(self, a0=None, T: int = 0):
        """fosmexpansion (mstm-scatprops-26.f90:1820-1951): s00, s02, sp22,
        sm22 as (2T+1, 4, 4) arrays [w, j-1, i-1] = s(i, j, w), with the
        reference's sign patch and 1/(2 m_medium) scale."""
        Tu = int(T) if a0 is not None else int(self.stats()["obs_T"])
        W = 2 * Tu + 1
        outs = [np.zeros((W, 4, 4)) for _ in range(4)]
        a = None if a0 is None else np.ascontiguousarray(a0, np.complex128)
        self._chk(self.L.mstm_hip_fosmexpansion(self.h, Tu, None if a is None else _d(a.view(np.float64)),
                                                *[_d(o) for o in outs]), "fosmexpansion")
        return tuple(outs)

    def observables(self, amnp=None, alpha: float = 0.0, eps: float = 1e-5, maxt: int = 100, copy: bool = False):
        """The single-origin post-processing of a solved cluster on the GPU:
        translation orders, merge, scattering matrix on the default grid
        (theta = -180..180 deg, scattering_angles) and g."""
        T, tord = self.translation_orders(None, eps, maxt)
        a0 = self.merge_to_common_origin(amnp, copy=copy)
        th, ct, phi = scattering_angles(alpha)
        sm = self.scattering_matrix(ct, phi)
        s00, g = self.asymmetry()
        return dict(T=T, torder=tord, amnp0=a0, theta=th, scatmat=sm, s00=s00, g=g)

    def direct_solver(self, pnp, initialize: bool = True, niter: int = 3):
        """direct_solver (mstm-solver-8.f90:571-642): (anp, solution_error)
        with anp = A^{-1} pnp by the LU factors of the dense interaction
        matrix A = I - T H, then niter refinement steps."""
        p = np.ascontiguousarray(pnp, np.complex128)
        out = np.zeros_like(p)
        e = np.zeros(1)
        self._chk(self.L.mstm_hip_direct_solver(self.h, _d(p.view(np.float64)), _d(out.view(np.float64)),
                                                int(bool(initialize)), int(niter), _d(e)), "direct_solver")
        return out, float(e[0])

    def tmatrix_solution(self, eps: float = 1e-6, niter: int = 10000, conveps: float = 1e-6, solver: int = 0,
                         teps: float = 1e-5, maxt: int = 100):
        """tmatrix_solution (mstm-solver-8.f90:22-332) on the GPU: the cluster
        T-matrix about the mean sphere position (translation orders from
        tranordertest at teps, as main_calling_program sets them).  Returns
        dict(T = order reached, tmat = the T-matrix file body in its write
        order, sqeff (nsphere, 3), qteff (3), qlog (T, 4): cumulative max
        iterations, Q_ext, Q_abs and the change of Q_ext per order; maxiter,
        maxerr, istat).  solver: SOLVE_PLAIN (the reference's FP64 iterates)
        or SOLVE_MIXED."""
        T, _ = self.translation_orders(None, teps, maxt)
        ntm = sum(4 * (2 * l + 1) * l * (l + 2) for l in range(1, T + 1))
        tm = np.zeros(ntm, np.complex128)
        sq = np.zeros((self.nsphere, 3))
        qt = np.zeros(3)
        ql = np.zeros((T, 4))
        tr = np.zeros(1, np.int32)
        mi = np.zeros(1, np.int32)
        me = np.zeros(1)
        ist = np.zeros(1, np.int32)
        self._chk(self.L.mstm_hip_tmatrix_solution(self.h, float(eps), int(niter), float(conveps), int(solver),
                                                   _d(tm.view(np.float64)), ntm, _d(sq), _d(qt), _d(ql), _i(tr),
                                                   _i(mi), _d(me), _i(ist)), "tmatrix_solution")
        Tr = int(tr[0])
        nused = sum(4 * (2 * l + 1) * l * (l + 2) for l in range(1, Tr + 1))
        return dict(T=Tr, T_in=T, tmat=tm[:nused], sqeff=sq, qteff=qt, qlog=ql[:Tr], maxiter=int(mi[0]),
                    maxerr=float(me[0]), istat=int(ist[0]))

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

    def set_refinement_cycles(self, cycles: int):
        """Largest number of FP64 refinement cycles of the mixed solvers
        (default 6); a polarization still above eps then falls back to the
        plain FP64 solver from b (0: at once)."""
        self._chk(self.L.mstm_hip_set_refinement_cycles(self.h, int(cycles)), "set_refinement_cycles")

    def set_graphs(self, on=True):
        """Replay small-cluster BiCG iterations from HIP graphs (default) or
        issue the same launches one by one (same results)."""
        self._chk(self.L.mstm_hip_set_graphs(self.h, int(bool(on))), "set_graphs")

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
