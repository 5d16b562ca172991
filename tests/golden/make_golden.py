"""Write the golden fixtures of tests/golden/ from the reference MSTM code.

Runs oracle/_ref/ref_harness (the unmodified reference modules compiled by
oracle/Makefile) on small clusters and stores inputs and full-precision
outputs as .npz.  Needs the reference tree (build container only); the
fixtures themselves are data and travel with the repository.

Cases
  cfg1_direct   20 spheres x=2, m=1.5+0.01i, seeded RSA in a sphere (fv 0.10):
                sphereinteraction on two seeded RHS, bare H, fixedorsoln + Q
  dense_l2      12 spheres x=0.1 (order 2 -> the reference's dense
                gentranmatrix path, translation_switch_order = 3)
  fft_cyl250    the first 250 spheres of the reference fixture
                examples/random_in_cylinder_1000.pos (unit spheres,
                m = 1.6+0.0123i as in the manual's example 6), FFT operator
                with automatic cell size and node order; plus the direct H.
                The FFT solve replays fixedorsoln with a 2-RHS initial call
                (see oracle/ref_harness.f90 for the reference allocation bug)
  frac_agg      examples/frac_agg.pos (100 touching monomers) scaled to
                x = 0.5, m = 1.75+0.4i: matvec + solve
  sparse100     100 spheres x=1, m=1.5+0.01i, seeded RSA in a sphere at
                fv 0.002 (common-origin order 49): solve + observables only
  poly30        30 spheres of mixed size (x = 0.4 .. 2.4, orders 2..6) on
                seeded RSA centres (placed at x = 2.4, fv 0.10, then shrunk,
                so they stay disjoint), m = 1.5+0.02i: direct matvec, bare H,
                solve, observables; and the same cluster through the FFT
                operator (poly30_fft)
  tmat5         5 spheres x=1, m=1.5+0.01i, seeded RSA in a sphere (fv 0.10):
                tmatrix_solution with the reference defaults (solution and
                convergence eps 1e-6), then ranorientscatmatrix on its file
                (orientation-averaged scattering-matrix expansion sm, smcf)
  tmat_poly6    6 spheres of mixed size (x = 0.3 .. 1.6 evenly, orders 3..5) on seeded
                RSA centres (placed at x = 1.6, fv 0.10), m = 1.4+0.05i:
                tmatrix_solution at solution eps 1e-8, convergence eps 1e-7
  lu_cfg1, lu_poly30  the clusters of cfg1_direct and poly30 through the
                direct (LU) branch of fixedorsoln (solution_method 'd'):
                solution and efficiencies
  cfg2_cut250   the 250 lowest spheres of the 1000-sphere x = 3 cylinder
                workload (seed 2, fv 0.25, m = 1.6+0.02i): FFT operator
                (grid 8x8x5, node order 9), direct H, FFT solve
  cfg4_cut300   the first 300 monomers of the 2000-monomer fractal workload
                (seed 4, x = 0.5, m = 1.75+0.4i): FFT operator (node order
                4), FFT solve at the workload's first orientation (from above)
  frac_agg_fft  examples/frac_agg.pos at x = 0.5 through the FFT operator
  boundary_cfg1 the cluster of cfg1_direct through fixedorsoln's optional
                branches: order of scattering (max_iterations -40, converged;
                -5, run out), max_iterations 0, an excitation radius (0.8 of
                the largest distance from the mean position, oblique
                incidence); and node_selection with the automatic cell size,
                a specified d_cell (2x), a target box widened by 3 and both
Every solve also runs the single-origin post-processing (translation orders,
merge_to_common_origin, scatteringmatrix on the default grid,
fosmexpansion); those outputs are stored under obs_* keys.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mstm_amd import geometry as G  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
REF_EXAMPLES = "/root/reference/examples"


def rhs(neqns, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(2, neqns)) + 1j * rng.normal(size=(2, neqns))


def case(name, pos, radius, ri, fft=False, solve=True, seed=0, observe=True, **incidence):
    ri = np.asarray(ri, np.complex128)
    setup = O.run_reference(pos, radius, ri, "setup")
    neqns = int(setup["neqns"])
    ain = rhs(neqns, seed)
    res = dict(pos=pos, radius=radius, ri=ri, ain=ain, order=setup["order"], offset=setup["offset"],
               anmie=setup["anmie"])
    mv = O.run_reference(pos, radius, ri, "matvec", ain=ain, use_fft=fft)
    res["aout"] = mv["aout"].reshape(2, neqns)
    if fft:
        res["cell_dim"] = np.array(mv["cell_dim"], np.int32)
        res["d_cell"] = mv["d_cell"]
        res["node_order"] = int(mv["node_order"])
        dm = O.run_reference(pos, radius, ri, "matvec", ain=ain, use_fft=False)
        res["aout_direct"] = dm["aout"].reshape(2, neqns)
        res["hout"] = dm["hout"].reshape(2, neqns)
    else:
        res["hout"] = mv["hout"].reshape(2, neqns)
    if solve:
        res.update(solve_case(pos, radius, ri, neqns, fft, observe, **incidence))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
    print(name, "neqns", neqns, "orders", sorted(set(res["order"].tolist())),
          "grid", (res.get("cell_dim"), res.get("node_order")), "iterations", res.get("iterations"))


def solve_case(pos, radius, ri, neqns, fft=False, observe=True, **incidence):
    """incidence: alpha, sinc, idir of fixedorsoln (default +z from below)."""
    sv = O.run_reference(pos, radius, ri, "solve", use_fft=fft, observe=observe, **incidence)
    out = dict(amnp=sv["amnp"].reshape(2, neqns), qeff=sv["qeff"].reshape(-1, 3, 3), qtot=sv["qtot"].reshape(3, 3),
               iterations=int(sv["iterations"]), maxerr=sv["maxerr"], istat=int(sv["istat"]),
               **{k: v for k, v in incidence.items()})
    if not observe:
        return out
    T = int(sv["t_matrix_order"])
    return dict(**out, obs_T=T, obs_torder=sv["torder"], obs_origin=np.array(sv["cluster_origin"]),
                obs_amnp0=sv["amnp0"].reshape(2, 2, T * (T + 2)), obs_scatmat=sv["scatmat"].reshape(361, 16),
                **{"obs_" + k: sv[k].reshape(2 * T + 1, 4, 4) for k in ("s00", "s02", "sp22", "sm22")})


def observe_case(name, pos, radius, ri):
    ri = np.asarray(ri, np.complex128)
    setup = O.run_reference(pos, radius, ri, "setup")
    neqns = int(setup["neqns"])
    res = dict(pos=pos, radius=radius, ri=ri, order=setup["order"], offset=setup["offset"])
    res.update(solve_case(pos, radius, ri, neqns))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
    print(name, "neqns", neqns, "T", res["obs_T"])


def lu_case(name, pos, radius, ri):
    ri = np.asarray(ri, np.complex128)
    r = O.run_reference(pos, radius, ri, "solve", method="'d'")
    ne = int(r["neqns"])
    res = dict(pos=pos, radius=radius, ri=ri, order=r["order"], amnp=r["amnp"].reshape(2, ne),
               qeff=r["qeff"].reshape(-1, 3, 3), qtot=r["qtot"].reshape(3, 3))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
    print(name, "neqns", ne)


def tmatrix_case(name, pos, radius, ri, eps, conveps, niter=10000):
    ri = np.asarray(ri, np.complex128)
    r = O.run_reference(pos, radius, ri, "tmatrix", eps=eps, niter=niter, conveps=conveps, ranorient=True)
    T = int(r["t_matrix_order"])
    res = dict(pos=pos, radius=radius, ri=ri, order=r["order"], eps=eps, conveps=conveps, niter=niter,
               rsm=r["rsm"].reshape(2 * T + 1, 4, 4), rsmcf=r["rsmcf"].reshape(2 * T + 1, 4, 4),
               torder=r["torder"], origin=np.array(r["cluster_origin"]), vol_radius=r["vol_radius"],
               T_in=int(r["t_matrix_order_in"]), T=int(r["t_matrix_order"]), istat=int(r["istat"]),
               tmat_header=r["tmat_header"], tmat=r["tmat"], sqeff=r["sqeff"].reshape(-1, 3), tmlog=r["tmlog"])
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
    print(name, "T", res["T_in"], "->", res["T"], "orders", sorted(set(res["order"].tolist())), "tmat", res["tmat"].size)


def excited_mask(pos, frac):
    """excited_spheres of an excitation radius frac * (largest distance from
    the mean position), as the harness forms it (excite_frac)."""
    d = np.sqrt(((pos - pos.mean(axis=0)) ** 2).sum(axis=1))
    return d <= frac * d.max()


def boundary_case(name, pos, radius, ri):
    """fixedorsoln's optional branches and node_selection's optional
    arguments through the reference: the order-of-scattering series
    (max_iterations < 0; converged and run out), no solve (0), an excitation
    radius, and the node grid with a specified cell size and a wider target
    box."""
    ri = np.asarray(ri, np.complex128)
    res = dict(pos=pos, radius=radius, ri=ri)
    for tag, kw in (("oos", dict(niter=-40, eps=1e-8)), ("oos_short", dict(niter=-5, eps=1e-8)),
                    ("noiter", dict(niter=0)), ("excited", dict(excite_frac=0.8, alpha=0.3, sinc=0.4))):
        r = O.run_reference(pos, radius, ri, "solve", **kw)
        ne = int(r["neqns"])
        res.update({f"{tag}_amnp": r["amnp"].reshape(2, ne), f"{tag}_qeff": r["qeff"].reshape(-1, 3, 3),
                    f"{tag}_iterations": int(r["iterations"]), f"{tag}_maxerr": float(r["maxerr"]),
                    f"{tag}_istat": int(r["istat"])})
        if "excited" in r:
            res["excited_count"] = int(r["excited"])
    auto = O.run_reference(pos, radius, ri, "setup", use_fft=True)
    res.update(grid_auto_cell_dim=np.array(auto["cell_dim"], np.int32), grid_auto_d_cell=float(auto["d_cell"]),
               grid_auto_fva=float(auto["cell_volume_fraction"]))
    dsp = 2.0 * float(auto["d_cell"])
    for tag, kw in (("dspec", dict(dspec=True, dcell_in=dsp)), ("box", dict(tbox_pad=3.0)),
                    ("box_dspec", dict(tbox_pad=3.0, dspec=True, dcell_in=dsp))):
        g = O.run_reference(pos, radius, ri, "setup", use_fft=True, **kw)
        res.update({f"grid_{tag}_cell_dim": np.array(g["cell_dim"], np.int32), f"grid_{tag}_d_cell": float(g["d_cell"]),
                    f"grid_{tag}_node_order": int(g["node_order"]),
                    f"grid_{tag}_fva": float(g["cell_volume_fraction"])})
    res["grid_dspec_in"] = dsp
    res["grid_box_pad"] = 3.0
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **res)
    print(name, {k: v for k, v in res.items() if np.ndim(v) == 0})


def config_cuts():
    """Pieces of the FFT workloads at their own geometry and optics: the 250
    lowest spheres (a slab at the full volume fraction) of the 1000-sphere
    x = 3 cylinder (seed 2, m = 1.6+0.02i), and the first 300 monomers (a
    connected sub-aggregate) of the 2000-monomer fractal (seed 4, x = 0.5,
    m = 1.75+0.4i) solved for its first orientation (incidence from above);
    the reference's own aggregate file through the FFT operator."""
    cl = G.rsa_in_cylinder(1000, 3.0, 0.25, seed=2, ri=1.6 + 0.02j)
    idx = np.sort(np.argsort(cl.pos[:, 2])[:250])
    case("cfg2_cut250", cl.pos[idx], cl.radius[idx], cl.ri[idx], fft=True, seed=17, observe=False)
    agg = G.fractal_aggregate(2000, 0.5, 1.8, 1.3, seed=4, ri=1.75 + 0.4j)
    rng = np.random.default_rng(4)
    beta, alpha = float(np.arccos(1 - 2 * rng.random())), float(2 * np.pi * rng.random())
    case("cfg4_cut300", agg.pos[:300], agg.radius[:300], agg.ri[:300], fft=True, seed=18, observe=False,
         alpha=alpha, sinc=float(np.sin(beta)), idir=1 if np.cos(beta) >= 0 else 2)
    fa = G.read_pos(os.path.join(REF_EXAMPLES, "frac_agg.pos"))
    case("frac_agg_fft", fa.pos * 0.5, fa.radius * 0.5, np.full(fa.n, 1.75 + 0.4j), fft=True, seed=19,
         observe=False)


def boundary_only():
    c = G.rsa_in_sphere(20, 2.0, 0.10, seed=1, ri=1.5 + 0.01j)
    boundary_case("boundary_cfg1", c.pos, c.radius, c.ri)


def main():
    c = G.rsa_in_sphere(20, 2.0, 0.10, seed=1, ri=1.5 + 0.01j)
    case("cfg1_direct", c.pos, c.radius, c.ri, seed=11)
    c = G.rsa_in_sphere(12, 0.1, 0.10, seed=3, ri=1.75 + 0.4j)
    case("dense_l2", c.pos, c.radius, c.ri, seed=12)
    cyl = G.read_pos(os.path.join(REF_EXAMPLES, "random_in_cylinder_1000.pos"))
    n = 250
    case("fft_cyl250", cyl.pos[:n], cyl.radius[:n], np.full(n, 1.6 + 0.0123j), fft=True, seed=13)
    agg = G.read_pos(os.path.join(REF_EXAMPLES, "frac_agg.pos"))
    case("frac_agg", agg.pos * 0.5, agg.radius * 0.5, np.full(agg.n, 1.75 + 0.4j), seed=14)
    c = G.rsa_in_sphere(100, 1.0, 0.002, seed=7, ri=1.5 + 0.01j)
    observe_case("sparse100", c.pos, c.radius, c.ri)
    c = G.rsa_in_sphere(30, 2.4, 0.10, seed=8, ri=1.5 + 0.02j)
    radius = np.round(np.random.default_rng(8).uniform(0.4, 2.4, c.n), 6)
    case("poly30", c.pos, radius, c.ri, seed=15)
    case("poly30_fft", c.pos, radius, c.ri, fft=True, seed=16)
    c = G.rsa_in_sphere(5, 1.0, 0.10, seed=21, ri=1.5 + 0.01j)
    tmatrix_case("tmat5", c.pos, c.radius, c.ri, eps=1e-6, conveps=1e-6)
    c = G.rsa_in_sphere(6, 1.6, 0.10, seed=24, ri=1.4 + 0.05j)
    radius = np.linspace(0.3, 1.6, c.n)
    tmatrix_case("tmat_poly6", c.pos, radius, c.ri, eps=1e-8, conveps=1e-7)
    lu_only()
    boundary_only()
    config_cuts()


def lu_only():
    c = G.rsa_in_sphere(20, 2.0, 0.10, seed=1, ri=1.5 + 0.01j)
    lu_case("lu_cfg1", c.pos, c.radius, c.ri)
    c = G.rsa_in_sphere(30, 2.4, 0.10, seed=8, ri=1.5 + 0.02j)
    radius = np.round(np.random.default_rng(8).uniform(0.4, 2.4, c.n), 6)
    lu_case("lu_poly30", c.pos, radius, c.ri)


def tmatrix_only():
    c = G.rsa_in_sphere(5, 1.0, 0.10, seed=21, ri=1.5 + 0.01j)
    tmatrix_case("tmat5", c.pos, c.radius, c.ri, eps=1e-6, conveps=1e-6)
    c = G.rsa_in_sphere(6, 1.6, 0.10, seed=24, ri=1.4 + 0.05j)
    radius = np.linspace(0.3, 1.6, c.n)
    tmatrix_case("tmat_poly6", c.pos, radius, c.ri, eps=1e-8, conveps=1e-7)


if __name__ == "__main__":
    {"tmatrix": tmatrix_only, "lu": lu_only, "boundary": boundary_only, "cuts": config_cuts}.get(sys.argv[1] if len(sys.argv) > 1 else "", main)()
