"""ctypes view of the C oracle (oracle/_ref/liboracle.so) and a runner for the
reference dump harness (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "liboracle.so")
HARNESS = os.path.join(REF_DIR, "ref_harness")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "lib"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        L.orc_cluster_new.restype = vp
        L.orc_cluster_new.argtypes = [C.c_int, dp, dp, vp, C.c_double, C.c_double, C.c_int]
        L.orc_cluster_free.argtypes = [vp]
        L.orc_cluster_neqns.argtypes = [vp]
        L.orc_cluster_orders.argtypes = [vp, ip, ip]
        L.orc_cluster_mie.argtypes = [vp, vp]
        L.orc_mie_mult.argtypes = [vp, C.c_int, C.c_int, vp, vp]
        L.orc_external_to_external.argtypes = [vp, C.c_int, ip, vp, vp]
        L.orc_sphereinteraction.argtypes = [vp, C.c_int, ip, ip, C.c_int, vp, vp]
        L.orc_planewave.argtypes = [vp, C.c_double, C.c_double, C.c_int, vp]
        L.orc_cbicg.argtypes = [vp, C.c_int, C.c_int, C.c_double, vp, vp, dp]
        L.orc_qeff.argtypes = [vp, vp, vp, dp]
        L.orc_qtot.argtypes = [vp, dp, dp]
        L.orc_fixedorsoln.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, vp, dp, dp,
                                      ip]
        L.orc_node_selection.argtypes = [vp, C.c_double, C.c_int, C.c_int, ip, dp, ip]
        L.orc_node_selection_ex.argtypes = [vp, dp, dp, dp, C.c_int, dp, C.c_int, C.c_int, ip, ip]
        L.orc_planewave_ex.argtypes = [vp, C.c_double, C.c_double, C.c_int, ip, vp]
        L.orc_fixedorsoln_ex.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, ip, vp,
                                         dp, dp, ip]
        L.orc_fft_external_to_external.argtypes = [vp, C.c_int, ip, vp, vp]
        L.orc_local_pairs.argtypes = [vp, ip]
        L.orc_rotcoef.argtypes = [C.c_double, C.c_int, C.c_int, dp]
        L.orc_normalizedlegendre.argtypes = [C.c_double, C.c_int, C.c_int, dp]
        L.orc_atcdim.argtypes = [C.c_int, C.c_int]
        L.orc_moffset.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_axialtrancoef.argtypes = [C.c_int, C.c_double, vp, C.c_int, C.c_int, vp]
        L.orc_gentranmatrix.argtypes = [C.c_int, C.c_int, dp, vp, C.c_int, vp]
        L.orc_riccati_psi.argtypes = [C.c_int, C.c_double, dp]
        L.orc_riccati_xi.argtypes = [C.c_int, C.c_double, vp]
        L.orc_logderiv.argtypes = [C.c_int, C.c_double, C.c_double, vp]
        L.orc_pair_translate.argtypes = [C.c_int, dp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
        L.orc_mie_order.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_mie_coef.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, vp, dp, dp]
        L.orc_tranorder.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double]
        L.orc_translation_orders.argtypes = [vp, dp, C.c_double, C.c_int, ip]
        L.orc_merge.argtypes = [vp, C.c_int, ip, dp, C.c_int, vp, vp]
        L.orc_scatmat.argtypes = [C.c_int, vp, C.c_double, C.c_double, C.c_double, dp, vp]
        L.orc_fosm.argtypes = [C.c_int, vp, C.c_double, dp, dp, dp, dp]
        L.orc_distribute.argtypes = [vp, C.c_int, dp, vp, vp]
        L.orc_ranorient.argtypes = [C.c_int, vp, C.c_double, C.c_double, dp, dp]
        L.orc_tmatrix.restype = C.c_int
        L.orc_tmatrix.argtypes = [vp, C.c_int, ip, dp, C.c_double, C.c_double, C.c_int, C.c_double, vp, dp, dp, dp,
                                  ip, dp, ip]
        L.orc_cg.restype = C.c_double
        L.orc_cg.argtypes = [C.c_int] * 5
        L.orc_bcof.restype = C.c_double
        L.orc_bcof.argtypes = [C.c_int, C.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _z2(z):
    return (C.c_double * 2)(z.real, z.imag)


class OracleCluster:
    """Cluster state of the C oracle (spheredata + mie arrays)."""

    def __init__(self, pos, radius, ri, ri_medium=1.0, mie_eps=1e-6, switch_order=3):
        L = lib()
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.ri = np.ascontiguousarray(ri, dtype=np.complex128)
        self.n = len(self.radius)
        self.h = L.orc_cluster_new(self.n, _d(self.pos), _d(self.radius), self.ri.ctypes.data, float(complex(ri_medium).real),
                                   mie_eps, switch_order)
        self.neqns = L.orc_cluster_neqns(self.h)
        self.order = np.zeros(self.n, np.int32)
        self.offset = np.zeros(self.n + 1, np.int32)
        L.orc_cluster_orders(self.h, _i(self.order), _i(self.offset))
        self.fft = None

    def __del__(self):
        try:
            lib().orc_cluster_free(self.h)
        except Exception:
            pass

    def mie(self):
        out = np.zeros(4 * int(self.order.sum()), np.complex128)
        lib().orc_cluster_mie(self.h, out.ctypes.data)
        return out

    def mie_mult(self, a, idir=1, inverse=False):
        a = np.ascontiguousarray(a, np.complex128)
        out = np.zeros_like(a)
        lib().orc_mie_mult(self.h, idir, int(inverse), a.ctypes.data, out.ctypes.data)
        return out

    def external_to_external(self, ain, contran):
        ain = np.ascontiguousarray(ain, np.complex128)
        nrhs = ain.shape[0]
        out = np.zeros_like(ain)
        ct = np.array(contran, np.int32)
        lib().orc_external_to_external(self.h, nrhs, _i(ct), ain.ctypes.data, out.ctypes.data)
        return out

    def sphereinteraction(self, ain, contran, miemult, use_fft=False):
        """ain: (nrhs, neqns).  Returns aout (nrhs, neqns)."""
        ain = np.ascontiguousarray(ain, np.complex128)
        nrhs = ain.shape[0]
        out = np.zeros_like(ain)
        ct = np.array(contran, np.int32)
        mm = np.array(miemult, np.int32)
        lib().orc_sphereinteraction(self.h, nrhs, _i(ct), _i(mm), int(use_fft), ain.ctypes.data, out.ctypes.data)
        return out

    def planewave(self, alpha=0.0, sinc=0.0, direction=1, excited=None):
        out = np.zeros((2, self.neqns), np.complex128)
        ex = None if excited is None else np.ascontiguousarray(excited, np.int32)
        lib().orc_planewave_ex(self.h, alpha, sinc, direction, None if ex is None else _i(ex), out.ctypes.data)
        return out

    def cbicg(self, pnp, anp0, niter=10000, eps=1e-6, use_fft=False):
        anp = np.ascontiguousarray(anp0, np.complex128).copy()
        pnp = np.ascontiguousarray(pnp, np.complex128)
        err = np.zeros(1)
        it = lib().orc_cbicg(self.h, int(use_fft), niter, eps, pnp.ctypes.data, anp.ctypes.data, _d(err))
        return anp, it, float(err[0])

    def fixedorsoln(self, alpha=0.0, sinc=0.0, eps=1e-6, niter=10000, use_fft=False, direction=1, excited=None):
        amnp = np.zeros((2, self.neqns), np.complex128)
        qeff = np.zeros((self.n, 3, 3))
        maxerr = np.zeros(1)
        maxiter = np.zeros(1, np.int32)
        ex = None if excited is None else np.ascontiguousarray(excited, np.int32)
        istat = lib().orc_fixedorsoln_ex(self.h, int(use_fft), alpha, sinc, direction, eps, niter,
                                         None if ex is None else _i(ex), amnp.ctypes.data, _d(qeff), _d(maxerr),
                                         _i(maxiter))
        return dict(amnp=amnp, qeff=qeff, maxerr=float(maxerr[0]), iterations=int(maxiter[0]), istat=istat)

    def qeff(self, amnp, pmnp):
        q = np.zeros((self.n, 3, 3))
        a = np.ascontiguousarray(amnp, np.complex128)  # held until the call returns
        p = np.ascontiguousarray(pmnp, np.complex128)
        lib().orc_qeff(self.h, a.ctypes.data, p.ctypes.data, _d(q))
        return q

    def qtot(self, qeff):
        q = np.zeros((3, 3))
        lib().orc_qtot(self.h, _d(np.ascontiguousarray(qeff)), _d(q))
        return q

    def node_selection(self, fva=0.0, node_order=-1, neighbor_model=2, target_min=None, target_max=None,
                       d_specified=False, d_cell=0.0):
        cd = np.zeros(3, np.int32)
        d = np.array([float(d_cell)])
        fv = np.array([float(fva)])
        no = np.zeros(1, np.int32)
        tmn = None if target_min is None else np.ascontiguousarray(target_min, np.float64)
        tmx = None if target_max is None else np.ascontiguousarray(target_max, np.float64)
        rc = lib().orc_node_selection_ex(self.h, _d(fv), None if tmn is None else _d(tmn),
                                         None if tmx is None else _d(tmx), int(bool(d_specified)), _d(d),
                                         node_order, neighbor_model, _i(cd), _i(no))
        if rc != 0:
            raise ValueError("node_selection: the target box has no extent along an axis")
        self.fft = dict(cell_dim=cd.copy(), d_cell=float(d[0]), node_order=int(no[0]), fva=float(fv[0]))
        return self.fft

    def translation_orders(self, origin=None, eps=1e-5, maxt=100):
        """(T, torder) about origin (default: mean sphere position)."""
        r0 = np.ascontiguousarray(self.pos.mean(axis=0) if origin is None else origin, np.float64)
        t = np.zeros(self.n, np.int32)
        T = lib().orc_translation_orders(self.h, _d(r0), eps, maxt, _i(t))
        return int(T), t

    def merge(self, T, torder, origin, amnp):
        """merge_to_common_origin of the (nrhs, neqns) solution: (nrhs, 2, T(T+2)) E/M."""
        a = np.ascontiguousarray(amnp, np.complex128)
        nrhs = a.shape[0]
        out = np.zeros((nrhs, 2, T * (T + 2)), np.complex128)
        t = np.ascontiguousarray(torder, np.int32)
        r0 = np.ascontiguousarray(origin, np.float64)
        lib().orc_merge(self.h, T, _i(t), _d(r0), nrhs, a.ctypes.data, out.ctypes.data)
        return out

    def interaction_matrix(self):
        """general_interaction_matrix with mie_mult (mstm-modules-33.f90:5854-5985):
        A = I - T H, column q = e_q - sphereinteraction(e_q) (dense, neqns^2)."""
        ne = self.neqns
        A = np.eye(ne, dtype=np.complex128)
        for q0 in range(0, ne, 256):
            q1 = min(ne, q0 + 256)
            e = np.zeros((q1 - q0, ne), np.complex128)
            e[np.arange(q1 - q0), np.arange(q0, q1)] = 1.0
            A[:, q0:q1] -= self.sphereinteraction(e, [0] * (q1 - q0), [1] * (q1 - q0)).T
        return A

    def direct_solve(self, b, A=None):
        """direct_solver (mstm-solver-8.f90:571-642) restated: the exact solution
        of A x = b for the columns of b (LAPACK LU of the dense A)."""
        A = self.interaction_matrix() if A is None else A
        return np.ascontiguousarray(np.linalg.solve(A, np.asarray(b, np.complex128).T).T)

    def distribute(self, l, origin, a0):
        """distribute_from_common_origin of one (2, l(l+2)) helicity column -> (neqns,)."""
        a = np.ascontiguousarray(a0, np.complex128)
        out = np.zeros(self.neqns, np.complex128)
        r0 = np.ascontiguousarray(origin, np.float64)
        lib().orc_distribute(self.h, l, _d(r0), a.ctypes.data, out.ctypes.data)
        return out

    def tmatrix(self, eps=1e-6, niter=10000, conveps=1e-6, teps=1e-5, T=None, torder=None, origin=None):
        """tmatrix_solution (mstm-solver-8.f90:22-332), serial iterative branch.

        Orders as main_calling_program sets them (origin = mean position,
        tranordertest at teps); returns dict with T (order reached), tmat (the
        T-matrix file body), sqeff (nsphere, 3), qteff, qlog (T, 4: cumulative
        max iterations, qext, qabs, converr), maxiter, maxerr, istat."""
        r0 = np.ascontiguousarray(self.pos.mean(axis=0) if origin is None else origin, np.float64)
        if torder is None:
            T0, torder = self.translation_orders(r0, teps)
            T = T0 if T is None else T
        t = np.ascontiguousarray(torder, np.int32)
        vr = 0.0  # vol_radius (mstm-input-37.f90:1205-1211), sequential sum, real(4) exponent
        for r in self.radius:
            vr += float(r) * float(r) * float(r)
        vr = vr ** float(np.float32(0.333333))
        ntm = sum(2 * (2 * l + 1) * 2 * l * (l + 2) for l in range(1, T + 1))
        tm = np.zeros(ntm, np.complex128)
        sq = np.zeros((self.n, 3))
        qt = np.zeros(3)
        ql = np.zeros((T, 4))
        mi = np.zeros(1, np.int32)
        me = np.zeros(1)
        ist = np.zeros(1, np.int32)
        Tr = lib().orc_tmatrix(self.h, T, _i(t), _d(r0), vr, eps, niter, conveps, tm.ctypes.data, _d(sq), _d(qt),
                               _d(ql), _i(mi), _d(me), _i(ist))
        nused = sum(2 * (2 * l + 1) * 2 * l * (l + 2) for l in range(1, Tr + 1))
        return dict(T=int(Tr), tmat=tm[:nused], sqeff=sq, qteff=qt, qlog=ql[:Tr], maxiter=int(mi[0]),
                    maxerr=float(me[0]), istat=int(ist[0]), vol_radius=vr, origin=r0, torder=t)

    def local_pairs(self):
        n = lib().orc_local_pairs(self.h, None)
        p = np.zeros((n, 2), np.int32)
        lib().orc_local_pairs(self.h, _i(p))
        return p


def ranorient(T, tmat, rimed=1.0, cbeam=0.0):
    """ranorientscatmatrix (mstm-scatprops-26.f90:1963-2530) on a T-matrix file
    body: (sm, smcf), each (2T+1, 4, 4) with [w][j-1][i-1] = Fortran sm(i,j,w)."""
    t = np.ascontiguousarray(tmat, np.complex128)
    sm = np.zeros((2 * T + 1, 4, 4))
    smcf = np.zeros((2 * T + 1, 4, 4))
    lib().orc_ranorient(int(T), t.ctypes.data, float(rimed), float(cbeam), _d(sm), _d(smcf))
    return sm, smcf


# the reference's pi: a real(4) literal widened to real(8) (mstm-modules-33.f90:18)
PI_R4 = float(np.float32(np.pi))


def scattering_angles(alpha=0.0):
    """Default scattering-matrix grid of a single-origin run without
    azimuthal averaging (mstm-input-37.f90:1339-1351, 3036-3046):
    theta = -180..180 deg in 1 deg steps; negative theta lies in the
    half-plane phi = alpha + pi.  Returns (theta_deg, cos theta, phi)."""
    th = np.arange(-180, 181)
    ct = np.cos(PI_R4 / 180.0 * th.astype(np.float64))
    ct[ct == 1.0] = 0.9999999
    ct[ct == -1.0] = -0.9999999
    phi = np.where(th < 0, alpha + PI_R4, alpha)
    return th, ct, phi


def scatmat(T, a0, ct, phi, rimed=1.0):
    """16 Mueller elements (Fortran sm(4,4) column-major) at each (ct, phi)."""
    a0 = np.ascontiguousarray(a0, np.complex128)
    ct = np.atleast_1d(ct)
    phi = np.broadcast_to(np.atleast_1d(phi), ct.shape)
    out = np.zeros((len(ct), 16))
    row = np.zeros(16)
    for k in range(len(ct)):
        lib().orc_scatmat(T, a0.ctypes.data, float(ct[k]), float(phi[k]), rimed, _d(row), None)
        out[k] = row
    return out


def fosm(T, a0, rimed=1.0):
    """fosmexpansion: s00, s02, sp22, sm22 as (2T+1, 4, 4) arrays [w, j-1, i-1] = s(i, j, w)."""
    a0 = np.ascontiguousarray(a0, np.complex128)
    outs = [np.zeros((2 * T + 1, 4, 4)) for _ in range(4)]
    lib().orc_fosm(T, a0.ctypes.data, rimed, *[_d(o) for o in outs])
    return outs


def observables(cl: "OracleCluster", amnp, alpha=0.0, eps=1e-5, maxt=100):
    """The single-origin post-processing of main_calling_program on a solved
    cluster: orders, merged coefficients, default-grid scattering matrix,
    s00 expansion and the asymmetry parameter g = s00(1,1,1)/(3 s00(1,1,0))."""
    r0 = cl.pos.mean(axis=0)
    T, tord = cl.translation_orders(r0, eps, maxt)
    a0 = cl.merge(T, tord, r0, amnp)
    th, ct, phi = scattering_angles(alpha)
    sm = scatmat(T, a0, ct, phi)
    s00 = fosm(T, a0)[0]
    return dict(T=T, torder=tord, origin=r0, amnp0=a0, theta=th, scatmat=sm, s00=s00,
                g=s00[1, 0, 0] / (3.0 * s00[0, 0, 0]))


def rotcoef(cbe, kmax, nmax):
    out = np.zeros(((nmax + 1) ** 2, 2 * kmax + 1))
    lib().orc_rotcoef(cbe, kmax, nmax, _d(out))
    return out  # out[n(n+1)+k, m+kmax] = d^n_{m k}


def axialtrancoef(itype, r, nmax, lmax, ri=1.0):
    L = lib()
    ri2 = np.array([ri, ri], np.complex128)
    out = np.zeros(L.orc_atcdim(nmax, lmax), np.complex128)
    L.orc_axialtrancoef(itype, r, ri2.ctypes.data, nmax, lmax, out.ctypes.data)
    return out


def gentranmatrix(nodr_s, nodr_t, xp, itype=3, ri=1.0):
    ri2 = np.array([ri, ri], np.complex128)
    xp = np.ascontiguousarray(xp, np.float64)
    out = np.zeros((2, nodr_s * (nodr_s + 2), nodr_t * (nodr_t + 2)), np.complex128)
    lib().orc_gentranmatrix(nodr_s, nodr_t, _d(xp), ri2.ctypes.data, itype, out.ctypes.data)
    return out  # out[p, source, target]


def pair_translate(itype, rtran, noa, nog, sop, top, ain, force_rot=False, ri=1.0):
    ri2 = np.array([ri, ri], np.complex128)
    g = np.zeros(2 * nog * (nog + 2), np.complex128)
    a = np.ascontiguousarray(ain, np.complex128)
    rt = np.ascontiguousarray(rtran, np.float64)
    lib().orc_pair_translate(itype, _d(rt), ri2.ctypes.data, noa, nog, int(sop), int(top), int(force_rot),
                             a.ctypes.data, g.ctypes.data)
    return g


def riccati_psi(n, x):
    out = np.zeros(n + 1)
    lib().orc_riccati_psi(n, x, _d(out))
    return out


def riccati_xi(n, x):
    out = np.zeros(n + 1, np.complex128)
    lib().orc_riccati_xi(n, x, out.ctypes.data)
    return out


def logderiv(n, z):
    out = np.zeros(n + 1, np.complex128)
    z = complex(z)
    lib().orc_logderiv(n, z.real, z.imag, out.ctypes.data)
    return out


def mie_order(x, m, mmed=1.0, qeps=1e-6):
    m = complex(m)
    return lib().orc_mie_order(x, m.real, m.imag, float(complex(mmed).real), qeps)


def mie_coef(x, m, nodr, mmed=1.0):
    out = np.zeros(4 * nodr, np.complex128)
    qe = np.zeros(1)
    qs = np.zeros(1)
    m = complex(m)
    lib().orc_mie_coef(x, m.real, m.imag, float(complex(mmed).real), nodr, out.ctypes.data, _d(qe), _d(qs))
    return out.reshape(nodr, 2, 2).transpose(0, 2, 1), float(qe[0]), float(qs[0])  # [n-1, p, q]


# ---------------------------------------------------------------- reference
def have_reference_harness() -> bool:
    return os.path.exists(HARNESS)


def run_reference(pos, radius, ri, mode, ain=None, outdir=None, timeout=None, **nml):
    """Run the reference dump harness on a cluster; returns dict of outputs."""
    if not have_reference_harness():
        raise FileNotFoundError(HARNESS)
    own = outdir is None
    outdir = outdir or tempfile.mkdtemp(prefix="mstmref_")
    posfile = os.path.join(outdir, "cluster.pos")
    with open(posfile, "w") as f:
        for p, r, m in zip(pos, radius, ri):
            f.write(f"{p[0]:.17e} {p[1]:.17e} {p[2]:.17e} {r:.17e} {m.real:.17e} {m.imag:.17e}\n")
    prefix = os.path.join(outdir, "out")
    items = dict(mode=f"'{mode}'", posfile=f"'{posfile}'", nsphere=len(radius), outprefix=f"'{prefix}'")
    if ain is not None:
        infile = os.path.join(outdir, "ain.bin")
        np.ascontiguousarray(ain, np.complex128).tofile(infile)
        items["infile"] = f"'{infile}'"
    for k, v in nml.items():
        if isinstance(v, bool):
            items[k] = ".true." if v else ".false."
        elif isinstance(v, float):
            items[k] = repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
        else:
            items[k] = str(v)
    jobfile = os.path.join(outdir, "job.nml")
    with open(jobfile, "w") as f:
        f.write("&job\n" + ",\n".join(f" {k}={v}" for k, v in items.items()) + "\n/\n")
    subprocess.run([HARNESS, jobfile], check=True, cwd=outdir, timeout=timeout,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    res = {}
    with open(prefix + "_info.txt") as f:
        for line in f:
            v = line.split()
            key, vals = v[0], v[1:]
            nums = [float(x.replace("D", "e")) for x in vals]
            res[key] = nums[0] if len(nums) == 1 else nums
    for name in os.listdir(outdir):
        if name.startswith("out_") and name.endswith(".bin"):
            key = name[4:-4]
            raw = open(os.path.join(outdir, name), "rb").read()
            if key in ("order", "offset", "torder"):
                res[key] = np.frombuffer(raw, np.int32).copy()
            elif key in ("qeff", "qtot", "rotcoef", "scatmat", "s00", "s02", "sp22", "sm22", "sqeff", "rsm", "rsmcf"):
                res[key] = np.frombuffer(raw, np.float64).copy()
            else:
                res[key] = np.frombuffer(raw, np.complex128).copy()
    tfile = prefix + "_tmat.dat"
    if os.path.exists(tfile):  # tmatrix_solution's T-matrix file (mstm-solver-8.f90:125-130, 243-256)
        with open(tfile) as f:
            head = f.readline().split()
            res["tmat_header"] = np.array([int(head[0]), int(head[1])], np.int32)
            vals = []
            for line in f:
                line = line.strip()
                if line.startswith("("):
                    re_, im_ = line.strip("()").split(",")
                    vals.append(complex(float(re_.replace("D", "e")), float(im_.replace("D", "e"))))
        res["tmat"] = np.array(vals, np.complex128)
    lfile = prefix + "_tmlog.txt"
    if os.path.exists(lfile):  # per-order lines: l, maxiter, qeff(1:2), converr, dqeffi(1,1), ...
        rows = []
        with open(lfile) as f:
            for line in f:
                v = line.split()
                if len(v) >= 6 and v[0].isdigit():
                    rows.append([float(x) for x in v[:6]])
        res["tmlog"] = np.array(rows)
    if own:
        import shutil
        shutil.rmtree(outdir, ignore_errors=True)
    return res
