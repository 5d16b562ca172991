"""Benchmark of the MSTM solution hot path on one MI355X.

Workload (the 5000-sphere dense cluster the headline metric is quoted on):
5000 spheres, size parameter x = 2.0, m = 1.31+0i, centres by seeded random
sequential addition in a sphere at volume fraction 0.30 (seed 5), plane wave
along +z, both polarizations, direct pair translation, BiCG to the
reference's stopping rule |r|^2/|b|^2 < 1e-6.

A step is one full fixed-orientation solve (fixedorsoln): incident-field
RHS, T-matrix multiply, solve for both polarizations, efficiencies.  The
learned + mixed path (the shipped per-sphere network's starting vector, FP32
translation tables and arithmetic inside FP64 iterative refinement) is timed
for K steps; the mixed path without the network and the plain FP64 GPU solve
are timed beside it, and a plain FP64 solve driven to |r|^2/|b|^2 < 1e-10 is
the reference for the Q_ext/Q_abs errors.  Translation tables are built once
per geometry (setup, reported separately, like the reference's
store_translation_matrix option).

Prints one JSON line.  --gpus N > 1 is not supported (one-GPU build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "s/solve to residual 1e-6 + matvec HBM GB/s; mixed/learned vs plain FP64 GPU"
HBM_PEAK_GBPS = 8000.0


def direct_sym(L: int, elem: int) -> bool:
    """The direct tiles store the symmetric table layout up to order 5 (FP32)
    / 4 (FP64) (mstm_amd/csrc/common.hpp, direct_sym)."""
    return L <= (5 if elem == 4 else 4)


def full_table_bytes_per_pair(L: int, elem: int) -> int:
    """Full rotation + axial table of one unordered pair (equal orders L):
    sum_{n<=L}(2n+1)^2 reals + 2 sum_{|m|<=L} (L-m1+1)^2 complex."""
    rot = sum((2 * n + 1) ** 2 for n in range(1, L + 1))
    ax = sum((L - max(1, abs(m)) + 1) ** 2 for m in range(-L, L + 1))
    return rot * elem + 2 * ax * 2 * elem


def sym_counts(L: int):
    """(rotation reals, axial complex) of the symmetric layout: blocks m >= 0
    only; half-sum/half-difference Wigner rows (n+1 values per degree for
    m = 0, 2n+1 for m > 0) and one axial block per mode (one for m = 0)."""
    rot = sum((n + 1) if m == 0 else (2 * n + 1) for m in range(0, L + 1) for n in range(max(1, m), L + 1))
    ax = sum((1 if m == 0 else 2) * (L - max(1, m) + 1) ** 2 for m in range(0, L + 1))
    return rot, ax


def table_bytes_per_pair(L: int, elem: int) -> int:
    """Table bytes of one unordered pair as the direct tiles store them."""
    if direct_sym(L, elem):
        rot, ax = sym_counts(L)
        return rot * elem + ax * 2 * elem
    return full_table_bytes_per_pair(L, elem)


def full_tile_flops_per_pair(L: int, nrhs: int = 2) -> int:
    """Flops of both directions of one pair for nrhs columns with full
    rotations: 16 (S_rot + S_ax) per direction and column (equal orders)."""
    s_rot = sum((2 * n + 1) ** 2 for n in range(1, L + 1))
    s_ax = sum((L - max(1, abs(m)) + 1) ** 2 for m in range(-L, L + 1))
    return 2 * nrhs * 16 * (s_rot + s_ax)


def tile_flops_per_pair(L: int, nrhs: int = 2, elem: int = 4) -> int:
    """Flops the direct tiles execute for both directions of one pair and
    nrhs columns: the symmetric rotations need R = sum of the stored row
    lengths multiply-adds per rotation instead of S_rot; axial work is
    unchanged: 16 (R + S_ax) per direction and column."""
    if not direct_sym(L, elem):
        return full_tile_flops_per_pair(L, nrhs)
    rot, _ = sym_counts(L)
    s_ax = sum((L - max(1, abs(m)) + 1) ** 2 for m in range(-L, L + 1))
    return 2 * nrhs * 16 * (rot + s_ax)


def cpu_baseline(cluster, iters_plain: int, budget_s: float = 20.0):
    """Reference MSTM (oracle/_ref/ref_harness, serial, 1 core) timed on a
    sample of the same cluster: sphereinteraction (2 RHS, direct pairs,
    tables regenerated per call as the reference's default) on the first
    n_s spheres; scaled by pair count to the full cluster and by the number
    of matvecs the FP64 BiCG solve needs (2 polarizations)."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return dict(value=None, unit="s/solve", cores=1, kind="reference", sample=f"unavailable: {e}")
    if not O.have_reference_harness():
        return dict(value=None, unit="s/solve", cores=1, kind="reference", sample="oracle/_ref not built")
    ns = 700
    n = cluster.n
    sub = slice(0, ns)
    L = 4
    ain = np.ones((2, ns * 2 * L * (L + 2)), np.complex128)
    t0 = time.time()
    r = O.run_reference(cluster.pos[sub], cluster.radius[sub], cluster.ri[sub], "matvec", ain=ain, timeout=600)
    wall = time.time() - t0
    per_pair = float(r["matvec_seconds"]) / (ns * (ns - 1) / 2)
    full_matvec = per_pair * n * (n - 1) / 2
    matvecs = 2 * (iters_plain + 1)
    return dict(value=full_matvec * matvecs, unit="s/solve", cores=1, kind="reference",
                sample=(f"reference sphereinteraction, 2 RHS, on the first {ns} spheres of the cluster "
                        f"({ns * (ns - 1) // 2} pairs, {r['matvec_seconds']:.2f} s cpu, {wall:.1f} s wall incl. setup); "
                        f"{per_pair * 1e6:.2f} us/pair-matvec x {n * (n - 1) // 2} pairs x {matvecs} matvecs "
                        f"(2 polarizations x (BiCG iterations of the FP64 solve + 1))"),
                s_per_matvec_full=full_matvec)


def _line(metric, value, unit, steps, warmup, ms, dtype, workload, extra, higher=False):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic (seeded)", "config": {"workload": workload}}
    out.update(extra)
    print(json.dumps(out))


def bench_batch(args):
    """256 clusters (N log-uniform 10-300, x 0.5-3, fv 0.05-0.30, m 1.33+0.001i),
    32 Fibonacci directions x 2 polarizations each, |r|^2/|b|^2 < 1e-6."""
    from concurrent.futures import ProcessPoolExecutor

    from mstm_amd import geometry as G
    from mstm_amd.batch import BatchSolver, random_clusters

    with ProcessPoolExecutor(8) as ex:  # geometry before the GPU is touched
        cl = list(ex.map(_batch_cluster, range(256), chunksize=8))
    b = BatchSolver(0)
    dirs = G.fibonacci_directions(32)
    t0 = time.perf_counter()
    b.set(cl, dirs)
    b.setup(tune=True)
    t_setup = time.perf_counter() - t0
    for _ in range(args.warmup):
        b.solve(eps=1e-6, niter=3000)
    times = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        r = b.solve(eps=1e-6, niter=3000)
        times.append(time.perf_counter() - t0)
        print(f"[bench] batch step {k}: {times[-1]:.3f} s", file=sys.stderr, flush=True)
    st = b.stats()
    nsolve = len(cl) * 64
    s_all = float(np.mean(times))
    q = b.qeff()
    tf = st["gemm32_flops"] / (st["gemm32_ms"] * 1e-3) / 1e12
    _line("s per batched solve (256 clusters x 64 right-hand sides) to residual 1e-6", round(s_all, 4), "s",
          args.steps, args.warmup, round(s_all * 1e3, 2), "c64 MFMA contraction + f64 refinement",
          "256 clusters N~logU[10,300] x~U[0.5,3] fv~U[0.05,0.3] m=1.33+0.001i, 32 Fibonacci directions x 2 pol",
          {"solves": nsolve, "s_per_solve": s_all / nsolve, "spheres": int(sum(c.n for c in cl)),
           "orders": np.bincount(b.order).tolist(), "setup_s": round(t_setup, 3),
           "build_ms": st["build_ms"], "autotune": {"grid": st["grid32"], "longest_first": bool(st["lpt"]),
                                                   "table": b.tune_table().tolist()},
           "iterations": {"max": int(r["iterations"].max()), "mean": float(r["iterations"].mean())},
           "max_err": float(r["err"].max()), "outer_steps": st["outer_steps"],
           "dense_operator_GB": {"fp32": st["hsize"] * 8 / 1e9, "fp64": st["hsize"] * 16 / 1e9},
           "qext_mean": float(q[:, :, 0, 0].mean()),
           "roofline": {"kernel": ("gemm32b_kernel" if st.get("gemm32_variant", 0) == 1 else "gemm32_kernel")
                        + " (grouped complex GEMM, v_mfma_f32_32x32x2f32)", "bound": "mfma",
                        "achieved": round(tf, 2), "peak": 157.3, "unit": "TFLOP/s", "frac": round(tf / 157.3, 4),
                        "traffic": None, "flops_per_launch": st["gemm32_flops"],
                        "algorithmic_bytes": st["gemm32_bytes"], "avg_launch_ms": st["gemm32_ms"]},
           "cpu_baseline": None})
    b.close()


def _batch_cluster(k):
    from mstm_amd.batch import random_clusters  # noqa: F401
    from mstm_amd import geometry as G

    rng = np.random.default_rng(3)
    ns = np.round(np.exp(rng.uniform(np.log(10), np.log(300), 256))).astype(int)
    xs = rng.uniform(0.5, 3.0, 256)
    fvs = rng.uniform(0.05, 0.30, 256)
    return G.rsa_in_sphere(int(ns[k]), float(xs[k]), float(fvs[k]), seed=3 * 100003 + k, ri=1.33 + 0.001j)


def bench_fft(args):
    """FFT node-grid operator: FP32 translation/FFT inside FP64 refinement vs
    the plain FP64 FFT solve, both checked against the FP64 direct solve."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import FP32, FP64, SOLVE_MIXED, SOLVE_PLAIN, HipContext, qtotcalc

    if args.workload == "fft-cylinder":
        cl = G.rsa_in_cylinder(1000, 3.0, 0.25, seed=2, ri=1.6 + 0.02j)
        dirs = [(0.0, 0.0)]
        name = "1000 spheres x=3 m=1.6+0.02i RSA cylinder (length 2 diameters) fv=0.25 seed=2, FFT, +z, 2 pol"
    else:
        cl = G.fractal_aggregate(2000, 0.5, 1.8, 1.3, seed=4, ri=1.75 + 0.4j)
        rng = np.random.default_rng(4)
        dirs = [(float(np.arccos(1 - 2 * rng.random())), float(2 * np.pi * rng.random())) for _ in range(8)]
        name = "2000-monomer aggregate Df=1.8 kf=1.3 x=0.5 m=1.75+0.4i seed=4, FFT, 8 orientations x 2 pol"
    ctx = HipContext(0)
    order, _ = ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    grid = ctx.node_selection()
    t0 = time.perf_counter()
    ctx.setup_fft(FP64)
    ctx.setup_fft(FP32)
    t_setup = time.perf_counter() - t0
    print(f"[bench] {cl.n} spheres, order {order.max()}, grid {grid}, FFT setup {t_setup:.2f} s", file=sys.stderr,
          flush=True)

    def solve_all(solver, fft):
        t0 = time.perf_counter()
        qs, its = [], []
        for beta, alpha in dirs:
            sd = 1 if np.cos(beta) >= 0 else 2
            r = ctx.fixedorsoln(alpha=alpha, sinc=float(np.sin(beta)), eps=1e-6, niter=10000, solver=solver,
                                fft_option=fft, direction=sd)
            qs.append(qtotcalc(cl.radius, r["qeff"]))
            its.append(r["iterations"])
        return (time.perf_counter() - t0) / len(dirs), np.array(qs), its

    for _ in range(args.warmup):
        solve_all(SOLVE_MIXED, True)
    ctx.clear(clear_tables=False)  # counters only
    times = []
    for _ in range(args.steps):
        t, q_mix, it_mix = solve_all(SOLVE_MIXED, True)
        times.append(t)
    t_mix = float(np.mean(times))
    st = ctx.stats()
    # the spectral contraction streams the node-kernel spectrum once per
    # two-column product: K (2 modes x nblk^2 x M complex FP32) + the
    # transformed node fields in and out (2 columns x 2 modes x nblk x M)
    M = 8 * int(np.prod(grid["cell_dim"]))
    LG = int(grid["node_order"])
    nblk = LG * (LG + 2)
    nc = max(1.0, st.get("fft_contract_launches_fp32", 0.0))
    t_c = st.get("fft_contract_ms_fp32", 0.0) / nc
    alg_c = 2 * nblk * nblk * M * 8 + 2 * (2 * 2 * nblk * M * 8)
    roof = None
    if t_c > 0:
        ach = alg_c / (t_c * 1e-3) / 1e9
        roof = dict(kernel="contract_kernel<float> (node-kernel spectrum x node fields, per frequency)", bound="hbm",
                    achieved=round(ach, 1), peak=HBM_PEAK_GBPS, unit="GB/s", frac=round(ach / HBM_PEAK_GBPS, 4),
                    traffic=None, algorithmic_bytes=alg_c, avg_launch_ms=round(t_c, 4), launches=int(nc),
                    flops_per_launch=2 * 2 * nblk * nblk * M * 8)
    conc = None
    if len(dirs) > 1 and not args.no_concurrent:
        conc = orientations_concurrent(cl, dirs, args.steps, q_mix)
    t_plain, q_plain, it_plain = solve_all(SOLVE_PLAIN, True)
    ctx.setup_direct(FP64)
    t_dir, q_dir, it_dir = solve_all(SOLVE_PLAIN, False)
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())  # noqa: E731
    _line("s/solve (both polarizations) to residual 1e-6, FFT operator", round(t_mix, 4), "s/solve", args.steps,
          args.warmup, round(t_mix * 1e3, 2), "c64 FFT/translation + f64 refinement", name,
          {"nsphere": cl.n, "order": int(order.max()),
           "node_grid": {k: (grid[k].tolist() if hasattr(grid[k], "tolist") else grid[k])
                         for k in ("cell_dim", "d_cell", "node_order", "fva")},
           "fft_setup_s": round(t_setup, 3), "iterations_mixed": it_mix,
           "plain_fft_fp64": {"s_per_solve": round(t_plain, 4), "iterations": it_plain},
           "direct_fp64": {"s_per_solve": round(t_dir, 4), "iterations": it_dir},
           "speedup_vs_plain_fft_fp64": round(t_plain / t_mix, 3),
           "q_error": {"mixed_fft_vs_plain_fft": rel(q_mix[:, 0, :2], q_plain[:, 0, :2]),
                       "mixed_fft_vs_direct": rel(q_mix[:, 0, :2], q_dir[:, 0, :2]),
                       "plain_fft_vs_direct": rel(q_plain[:, 0, :2], q_dir[:, 0, :2])},
           "orientations_concurrent": conc, "roofline": roof, "cpu_baseline": None})
    ctx.close()


def orientations_concurrent(cl, dirs, steps, q_serial):
    """The orientations at once: one context (own stream, FFT operator and
    Krylov vectors) and one host thread per orientation, each running the
    mixed FFT solve of its direction; the library calls release the GIL.
    Returns the time per orientation and the largest Q difference from the
    serial solves (the solves are the same, so it is 0)."""
    import threading

    from mstm_amd.solver import FP32, FP64, SOLVE_MIXED, HipContext, qtotcalc

    ctxs = []
    for _ in dirs:
        c = HipContext(0)
        c.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
        c.node_selection()
        c.setup_fft(FP64)
        c.setup_fft(FP32)
        ctxs.append(c)
    out = [None] * len(dirs)

    def run(k):
        beta, alpha = dirs[k]
        sd = 1 if np.cos(beta) >= 0 else 2
        out[k] = ctxs[k].fixedorsoln(alpha=alpha, sinc=float(np.sin(beta)), eps=1e-6, niter=10000,
                                     solver=SOLVE_MIXED, fft_option=True, direction=sd)

    def once():
        th = [threading.Thread(target=run, args=(k,)) for k in range(len(dirs))]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return (time.perf_counter() - t0) / len(dirs)

    once()  # warm-up
    ts = [once() for _ in range(max(1, steps))]
    q = np.array([qtotcalc(cl.radius, r["qeff"]) for r in out])
    d = float(np.abs(q[:, 0, :2] - q_serial[:, 0, :2]).max())
    for c in ctxs:
        c.close()
    return {"contexts": len(dirs), "s_per_orientation": round(float(np.mean(ts)), 4), "q_abs_diff_vs_serial": d}


def bench_tmatrix(args):
    """T-matrix solution (tmatrix_solution, mstm-solver-8.f90:22-332) of a
    200-sphere cluster (x = 1, m = 1.5+0.01i, seeded RSA in a sphere at fv 0.20)
    with the reference defaults (solution and convergence eps 1e-6): every
    order to convergence, 2(2l+1) unit-excitation solves per order.  The FP64
    solver with cbicg's iterates and the mixed FP32/FP64 solver are timed; the
    reference (oracle/_ref, serial) is timed on the order-1 columns of the
    same cluster and scaled by the number of solves."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import SOLVE_MIXED, SOLVE_PLAIN, HipContext

    cl = G.rsa_in_sphere(200, 1.0, 0.20, seed=41, ri=1.5 + 0.01j)
    ctx = HipContext(0)
    ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    if args.no_graphs:
        ctx.set_graphs(False)
    res = {}
    for name, solver in (("plain_fp64", SOLVE_PLAIN), ("mixed", SOLVE_MIXED)):
        for _ in range(args.warmup):
            ctx.tmatrix_solution(eps=1e-6, conveps=1e-6, solver=solver)
        times = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            r = ctx.tmatrix_solution(eps=1e-6, conveps=1e-6, solver=solver)
            times.append(time.perf_counter() - t0)
            print(f"[bench] tmatrix {name} step {k}: {times[-1]:.3f} s (order {r['T']})", file=sys.stderr, flush=True)
        st = ctx.stats()
        res[name] = dict(s=float(np.mean(times)), T=r["T"], T_in=r["T_in"], solves=int(st["tmatrix_solves"]),
                         maxiter=r["maxiter"], istat=r["istat"], qext=float(r["qlog"][-1, 1]),
                         qabs=float(r["qlog"][-1, 2]), tmat=r["tmat"])
    d = float(np.abs(res["mixed"]["tmat"] - res["plain_fp64"]["tmat"]).max() / np.abs(res["plain_fp64"]["tmat"]).max())
    nsolve = res["plain_fp64"]["solves"]
    cpu = None
    if not args.no_cpu_baseline:
        try:
            from oracle import oracle as O

            if O.have_reference_harness():
                t0 = time.time()
                rr = O.run_reference(cl.pos, cl.radius, cl.ri, "tmatrix", eps=1e-6, niter=10000, conveps=1e-6,
                                     tcap=1, timeout=600)
                wall = time.time() - t0
                per = float(rr["solve_seconds"]) / 6.0
                cpu = dict(value=per * nsolve, unit="s", cores=1, kind="reference",
                           sample=(f"reference tmatrix_solution (serial) on the order-1 columns of the same cluster: "
                                   f"6 solves in {rr['solve_seconds']:.2f} s ({wall:.1f} s wall incl. setup), "
                                   f"{per:.2f} s/solve x {nsolve} solves"))
        except Exception as e:  # pragma: no cover
            cpu = dict(value=None, unit="s", cores=1, kind="reference", sample=f"unavailable: {e}")
    best = min(("plain_fp64", "mixed"), key=lambda k: res[k]["s"])
    st = ctx.stats()
    _line("s per T-matrix solution (all orders to convergence, residual 1e-6)", round(res[best]["s"], 4), "s",
          args.steps, args.warmup, round(res[best]["s"] * 1e3, 2), "f64 (plain) / f32+f64 (mixed)",
          "200 spheres x=1.0 m=1.5+0.01i RSA sphere fv=0.20 seed=41, T-matrix to convergence eps 1e-6",
          {"solver": best, "order": res[best]["T"], "order_bound": res[best]["T_in"], "solves": nsolve,
           **{k: {kk: v for kk, v in res[k].items() if kk != "tmat"} for k in res},
           "mixed_vs_plain_tmat_rel": d, "fp64_products": int(st["matvecs_fp64"]),
           "fp32_products": int(st["matvecs_fp32"]), "cpu_baseline": cpu})
    ctx.close()


def bench_lu(args):
    """Direct (LU) solution of fixedorsoln (solution_method 'd') for a
    200-sphere cluster (x = 2, order 5, 14000 equations): dense A = I - T H
    from the FP64 operator, blocked LU with scaled partial pivoting, both
    polarizations back-substituted.  The factorization is timed against the
    FP64 vector peak (complex LU: 8 n^3 / 3 flops)."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import SOLVE_LU, SOLVE_PLAIN, HipContext

    cl = G.rsa_in_sphere(200, 2.0, 0.20, seed=17, ri=1.31 + 0.01j)
    ctx = HipContext(0)
    ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    times, st = [], None
    for k in range(args.warmup + args.steps):
        ctx.clear(clear_tables=True)
        t0 = time.perf_counter()
        r = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_LU)
        dt = time.perf_counter() - t0
        if k >= args.warmup:
            times.append(dt)
            st = ctx.stats()
        print(f"[bench] lu step {k}: {dt:.3f} s", file=sys.stderr, flush=True)
    ne = r["amnp"].shape[1]
    t0 = time.perf_counter()
    rp = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_PLAIN)
    t_plain = time.perf_counter() - t0
    d = float(np.abs(r["amnp"] - rp["amnp"]).max() / np.abs(rp["amnp"]).max())
    fl = 8.0 * ne ** 3 / 3.0
    tf = fl / (st["lu_factor_ms"] * 1e-3) / 1e12
    s = float(np.mean(times))
    _line("s per direct (LU) solution, both polarizations", round(s, 4), "s", args.steps, args.warmup,
          round(s * 1e3, 2), "f64", "200 spheres x=2.0 m=1.31+0.01i RSA sphere fv=0.20 seed=17, order 5, "
          f"{ne} equations, fixedorsoln direct branch",
          {"number_eqns": ne, "build_ms": st["lu_build_ms"], "factor_ms": st["lu_factor_ms"],
           "solve_ms": st["lu_solve_ms"], "plain_fp64_bicg_s": round(t_plain, 4),
           "lu_vs_bicg_1e-6_rel": d,
           "roofline": {"kernel": "lu_factor (panels + swaps + TRSM + complex FP64 GEMM on the matrix cores)",
                        "bound": "fp64 (matrix cores and vector: both 78.6 TF/s)",
                        "achieved": round(tf, 2), "peak": 78.6, "unit": "TFLOP/s", "frac": round(tf / 78.6, 4),
                        "traffic": None, "flops": fl},
           "cpu_baseline": None})
    ctx.close()


def small_cluster_solve(reps: int = 20):
    """Solve latency of the 20-sphere cluster (x = 2, m = 1.5+0.01i, seeded
    RSA in a sphere at fv 0.10, seed 1): plain FP64 BiCG with direct
    translation to |r|^2/|b|^2 < 1e-6, both polarizations; median of reps
    solves after two warm-up solves (tables built once, outside)."""
    from mstm_amd import geometry as G
    from mstm_amd.solver import FP64, HipContext

    c = G.rsa_in_sphere(20, 2.0, 0.10, seed=1, ri=1.5 + 0.01j)
    ctx = HipContext(0)
    try:
        order, _ = ctx.set_spheres(c.pos, c.radius, c.ri, 1.0, 1e-6)
        ctx.setup_direct(FP64)
        for _ in range(2):
            r = ctx.fixedorsoln(eps=1e-6, niter=10000)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            r = ctx.fixedorsoln(eps=1e-6, niter=10000)
            ts.append(time.perf_counter() - t0)
        return dict(workload="20 spheres x=2.0 m=1.5+0.01i RSA sphere fv=0.10 seed=1, direct, plain FP64 BiCG, 2 pol",
                    s_per_solve=round(float(np.median(ts)), 6), iterations=r["iterations"], maxerr=r["maxerr"],
                    order=int(order.max()), number_eqns=ctx.number_eqns, reps=reps)
    finally:
        ctx.close()


def observe(ctx):
    """Single-origin observables of the solution the last fixedorsoln left on the device."""
    ob = ctx.observables(None)
    ctx.fosmexpansion()
    st = ctx.stats()
    return dict(T=ob["T"], g=ob["g"], s11=ob["scatmat"][:, 0].copy(), s12=ob["scatmat"][:, 4].copy(),
                ms=dict(orders=round(st["obs_orders_ms"], 3), merge=round(st["obs_merge_ms"], 3),
                        scatmat=round(st["obs_scatmat_ms"], 3), g=round(st["obs_moments_ms"], 3),
                        fosmexpansion=round(st["obs_fosm_ms"], 3)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsphere", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tight", action="store_true", help="skip the tight FP64 reference solve")
    ap.add_argument("--no-concurrent", action="store_true",
                    help="fractal: skip the orientations-at-once run (for kernel tracing: its launches overlap)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="tmatrix: issue the small-cluster BiCG launches one by one instead of replaying HIP graphs "
                         "(for kernel tracing; same results)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB (rocprof FETCH/WRITE calibration)")
    ap.add_argument("--workload", default="direct5000",
                    choices=["direct5000", "fft-cylinder", "fractal", "batch", "tmatrix", "lu"],
                    help="direct5000 (default, the headline line); the others print their own line: "
                         "fft-cylinder = 1000 spheres x=3 m=1.6+0.02i in a cylinder at fv 0.25 (FFT path), "
                         "fractal = 2000-monomer aggregate x=0.5 m=1.75+0.4i (FFT path, 8 orientations), "
                         "batch = 256 clusters x 32 directions x 2 polarizations (batched MFMA solver), "
                         "tmatrix = T-matrix of a 200-sphere cluster x=1 (all orders to convergence), "
                         "lu = direct (LU) solution of a 200-sphere cluster x=2 (14000 equations)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        x = torch.ones(args.copy_probe * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        for _ in range(3):
            y = x.clone()
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 4, "launches": 3, "checksum": float(y[-1])}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload in ("fft-cylinder", "fractal"):
        return bench_fft(args)
    if args.workload == "tmatrix":
        return bench_tmatrix(args)
    if args.workload == "lu":
        return bench_lu(args)

    from mstm_amd import geometry as G
    from mstm_amd.solver import (FP32, FP64, FP64_FP32, SOLVE_FGMRES, SOLVE_LEARNED, SOLVE_MIXED, SOLVE_PLAIN,
                                 HipContext, qtotcalc)

    t0 = time.time()
    cl = G.rsa_in_sphere(args.nsphere, 2.0, 0.30, seed=5, ri=1.31 + 0.0j)
    t_geom = time.time() - t0
    ctx = HipContext(0)
    order, _ = ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    L = int(order.max())
    ne = ctx.number_eqns
    print(f"[bench] {cl.n} spheres, order {L}, neqns {ne}", file=sys.stderr, flush=True)
    t0 = time.time()
    ctx.setup_direct(FP64_FP32)  # both tables in one pass
    t_tables = time.time() - t0
    st0 = ctx.stats()
    print(f"[bench] tables built in {t_tables:.2f} s", file=sys.stderr, flush=True)

    meta = ctx.load_learned()

    # ---- learned + mixed path: warmup + timed steps --------------------------
    for _ in range(args.warmup):
        ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_LEARNED)
    ctx.clear(clear_tables=False)  # reset counters only
    times, res = [], None
    for k in range(args.steps):
        t0 = time.perf_counter()
        res = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_LEARNED)
        times.append(time.perf_counter() - t0)
        print(f"[bench] learned step {k}: {times[-1]:.3f} s, iters {res['iterations']}, err {res['maxerr']:.2e}",
              file=sys.stderr, flush=True)
    st_lig = ctx.stats()
    lig_s = float(np.mean(times))
    q_lig = qtotcalc(cl.radius, res["qeff"])
    ob_lig = observe(ctx)

    # ---- mixed path without the network --------------------------------------
    ctx.clear(clear_tables=False)
    mtimes = []
    for k in range(max(1, min(2, args.steps))):
        t0 = time.perf_counter()
        rm = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_MIXED)
        mtimes.append(time.perf_counter() - t0)
    st_mixed = ctx.stats()
    mixed_s = float(np.mean(mtimes))
    q_mixed = qtotcalc(cl.radius, rm["qeff"])
    ob_mixed = observe(ctx)
    print(f"[bench] mixed (no network): {mixed_s:.3f} s, iters {rm['iterations']}", file=sys.stderr, flush=True)

    # ---- flexible GMRES with the learned first preconditioner (once) ---------
    ctx.clear(clear_tables=False)
    t0 = time.perf_counter()
    rf = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_FGMRES)
    fg_s = time.perf_counter() - t0
    st_fg = ctx.stats()
    q_fg = qtotcalc(cl.radius, rf["qeff"])
    print(f"[bench] flexible GMRES: {fg_s:.3f} s, steps {st_fg['fgmres_steps']:.0f}", file=sys.stderr, flush=True)

    # ---- plain FP64 path (once) ----------------------------------------------
    plain = None
    if not args.no_plain:
        ctx.clear(clear_tables=False)
        t0 = time.perf_counter()
        rp = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_PLAIN)
        plain_s = time.perf_counter() - t0
        st_plain = ctx.stats()
        q_plain = qtotcalc(cl.radius, rp["qeff"])
        ob_plain = observe(ctx)
        n4 = st_plain.get("tile4_launches_fp64", 0)
        plain = dict(s_per_solve=plain_s, iterations_last_pol=rp["iterations"], maxerr=rp["maxerr"],
                     qext=q_plain[0, 0], qabs=q_plain[0, 1],
                     tile_columns=4 if n4 > 0 else 2,
                     tile_ms_avg=(st_plain["tile4_ms_fp64"] / n4 if n4 > 0 else
                                  st_plain["tile_ms_fp64"] / max(1, st_plain["tile_launches_fp64"])),
                     iters_fp64=st_plain["iters_fp64"])
        print(f"[bench] plain FP64 solve: {plain_s:.3f} s, iters {rp['iterations']}", file=sys.stderr, flush=True)

    # ---- tight FP64 reference for the efficiency errors ----------------------
    qerr = None
    if not args.no_tight:
        ctx.clear(clear_tables=False)
        rt = ctx.fixedorsoln(eps=1e-10, niter=10000, solver=SOLVE_PLAIN)
        q_ref = qtotcalc(cl.radius, rt["qeff"])
        ob_ref = observe(ctx)
        qerr = {}
        runs = (("learned", q_lig, ob_lig), ("mixed", q_mixed, ob_mixed)) + \
            ((("plain", q_plain, ob_plain),) if plain else ())
        qerr["fgmres"] = dict(qext=abs(q_fg[0, 0] - q_ref[0, 0]) / abs(q_ref[0, 0]))
        for name, q, ob in runs:
            qerr[name] = dict(qext=abs(q[0, 0] - q_ref[0, 0]) / abs(q_ref[0, 0]),
                              qsca=abs(q[0, 2] - q_ref[0, 2]) / abs(q_ref[0, 2]),
                              qabs_abs=abs(q[0, 1] - q_ref[0, 1]),
                              g_abs=abs(ob["g"] - ob_ref["g"]),
                              s11_max_rel=float(np.max(np.abs(ob["s11"] - ob_ref["s11"]) / np.abs(ob_ref["s11"]))),
                              s11_max_abs_over_peak=float(np.max(np.abs(ob["s11"] - ob_ref["s11"])) /
                                                          np.max(np.abs(ob_ref["s11"]))),
                              s12_max_rel_s11=float(np.max(np.abs(ob["s12"] - ob_ref["s12"]) / np.abs(ob_ref["s11"]))))
        qerr["reference"] = dict(eps=1e-10, qext=q_ref[0, 0], qabs=q_ref[0, 1], qsca=q_ref[0, 2], g=ob_ref["g"],
                                 s11_forward=ob_ref["s11"][180], s11_back=ob_ref["s11"][360],
                                 t_matrix_order=ob_ref["T"])
        qerr["note"] = ("g from the s00 moments of the merged single-origin expansion; S11/S12 on the default "
                        "grid theta = -180..180 deg (361 directions); errors against the FP64 solve at 1e-10. "
                        "s11_max_rel is per direction (largest in the deep minima of S11, set by the 1e-6 "
                        "residual: the plain FP64 solve shows it too); s11_max_abs_over_peak is relative to the "
                        "forward peak")
        print(f"[bench] tight FP64 reference done", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (FP32 tile kernel) ------------------
    npairs = cl.n * (cl.n - 1) // 2
    # both polarizations' BiCG products run as one four-column launch while
    # both are iterating; the roofline is quoted on that launch when present
    ncol = 4 if st_lig.get("tile4_launches_fp32", 0) > 0 else 2
    vec_bytes = ncol * 2 * ne * 8
    # algorithmic bytes of one product: the minimal per-pair table (rotation
    # rows sum_n (2n+1)^2 reals + both modes' axial blocks) at the stored
    # precision for every unordered pair, plus the vectors; the tiles store
    # and read the symmetric half of it (stored_bytes)
    alg32 = npairs * full_table_bytes_per_pair(L, 4) + vec_bytes
    st32 = npairs * table_bytes_per_pair(L, 4) + vec_bytes
    cols64 = plain["tile_columns"] if plain else 2
    alg64 = npairs * full_table_bytes_per_pair(L, 8) + cols64 * 2 * ne * 16
    if ncol == 4:
        t32 = st_lig["tile4_ms_fp32"] / st_lig["tile4_launches_fp32"]
    else:
        t32 = st_lig["tile_ms_fp32"] / max(1, st_lig["tile_launches_fp32"])
    ach32 = alg32 / (t32 * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(HERE, "profiles", "tile_fp32_cfg5_pmc.json")
    if os.path.exists(pmc):
        rec = json.load(open(pmc)).get("by_columns", {}).get(str(ncol))
        traffic = rec["hbm_bytes_per_launch"] if rec else None
    fl32 = npairs * full_tile_flops_per_pair(L, ncol)
    ex32 = npairs * tile_flops_per_pair(L, ncol, 4)
    roof = dict(kernel=f"direct_tile_kernel<float,{L},{ncol}>", bound="hbm", achieved=round(ach32, 1),
                peak=HBM_PEAK_GBPS, unit="GB/s", frac=round(ach32 / HBM_PEAK_GBPS, 4), traffic=traffic,
                algorithmic_bytes=alg32, columns=ncol, avg_launch_ms=round(t32, 4),
                stored_bytes=st32, stored_GBps=round(st32 / (t32 * 1e-3) / 1e9, 1),
                flops_per_launch=fl32, tflops=round(fl32 / (t32 * 1e-3) / 1e12, 2),
                executed_flops=ex32, executed_tflops=round(ex32 / (t32 * 1e-3) / 1e12, 2),
                flop_per_byte=round(fl32 / alg32, 2),
                note=("algorithmic bytes and flops per launch = the per-pair minimal table "
                      f"({full_table_bytes_per_pair(L, 4)} B at order {L} in FP32) and full-rotation flops "
                      "16 (S_rot + S_ax) per pair, direction and column, times the unordered pairs, plus the "
                      "vectors; the kernel stores and reads the symmetric half-table (stored_bytes: blocks "
                      "m >= 0, half-sum/half-difference Wigner rows, one axial block per mode), staged in LDS "
                      "once per tile step, and executes executed_flops; traffic is the measured HBM bytes; "
                      "packed FP32 is disabled, so the FP32 vector peak is 78.6 TF/s"))
    matvec = dict(fp32_tile_ms=round(t32, 4), fp32_columns=ncol, fp32_hbm_GBps=round(ach32, 1))
    if st_lig.get("tile_launches_fp32", 0) > 0:
        t2 = st_lig["tile_ms_fp32"] / st_lig["tile_launches_fp32"]
        a2 = npairs * full_table_bytes_per_pair(L, 4) + 2 * 2 * ne * 8
        matvec.update(fp32_2col_tile_ms=round(t2, 4), fp32_2col_hbm_GBps=round(a2 / (t2 * 1e-3) / 1e9, 1))
    if plain is not None:
        t64 = plain["tile_ms_avg"]
        matvec.update(fp64_tile_ms=round(t64, 4), fp64_columns=cols64,
                      fp64_hbm_GBps=round(alg64 / (t64 * 1e-3) / 1e9, 1),
                      fp64_frac=round(alg64 / (t64 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4))

    small = small_cluster_solve()
    print(f"[bench] 20-sphere solve: {small['s_per_solve'] * 1e3:.3f} ms", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cl, plain["iters_fp64"] // 2 if plain else 60)

    lsolves = max(1.0, st_lig["lig_solves"])
    out = {
        "metric": METRIC,
        "value": round(lig_s, 4),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(lig_s * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 translation (FP32) + f64 refinement",
        "data": "synthetic (seeded random sequential addition)",
        "config": {"workload": f"{cl.n} spheres x=2.0 m=1.31 RSA sphere fv=0.30 seed=5, direct translation, "
                               "plane wave +z, 2 pol, |r|^2/|b|^2<1e-6 (learned start + mixed FP32/FP64)",
                   "nsphere": cl.n, "order": L, "number_eqns": ne, "pairs": npairs},
        "time_to_solution_s": round(t_tables + lig_s, 4),
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(plain["s_per_solve"] / lig_s, 3) if plain else None,
        "learned": dict(iterations_last_pol=res["iterations"], maxerr=res["maxerr"],
                        fp64_rel_residual=float(np.sqrt(res["maxerr"])), qext=q_lig[0, 0],
                        qabs=q_lig[0, 1], iters_fp32=st_lig["iters_fp32"] / args.steps,
                        iters_fp64_matvecs=st_lig["matvecs_fp64"] / args.steps,
                        refinements=st_lig["refinements"] / args.steps,
                        inner_solver="gmres" if st_lig.get("inner_solver", 0) == 1 else "bicg",
                        fallback_rate=st_lig["lig_fallbacks"] / lsolves,
                        start_residual_learned=st_lig["lig_err0"], start_residual_default=st_lig["lig_errb"],
                        network_ms=st_lig["lig_ms"], network=dict(samples=meta.get("samples"),
                                                                 configs=meta.get("configs"),
                                                                 loss=meta.get("loss_final"))),
        "mixed": dict(s_per_solve=round(mixed_s, 4), iterations_last_pol=rm["iterations"], maxerr=rm["maxerr"],
                      fp64_rel_residual=float(np.sqrt(rm["maxerr"])),
                      qext=q_mixed[0, 0], qabs=q_mixed[0, 1], iters_fp32=st_mixed["iters_fp32"] / len(mtimes),
                      iters_fp64_matvecs=st_mixed["matvecs_fp64"] / len(mtimes),
                      refinements=st_mixed["refinements"] / len(mtimes),
                      speedup_vs_plain_fp64=round(plain["s_per_solve"] / mixed_s, 3) if plain else None),
        "fgmres": dict(s_per_solve=round(fg_s, 4), steps=st_fg["fgmres_steps"], iters_fp32=st_fg["iters_fp32"],
                       maxerr=rf["maxerr"], qext=q_fg[0, 0], fallbacks=st_fg["lig_fallbacks"]),
        "q_error_vs_tight_fp64": qerr,
        "observables": dict(t_matrix_order=ob_lig["T"], g=ob_lig["g"], ms=ob_lig["ms"],
                            note="post-solve, on the GPU: translation orders (host), merge to the common origin, "
                                 "scattering matrix at 361 directions, s00 moments (g), the full fosmexpansion"),
        "matvec": matvec,
        "small_cluster": small,
        "setup_s": dict(geometry=round(t_geom, 2), tables_fp64_fp32=round(t_tables, 3),
                        tables_one_pass_ms=st0["tables_ms_fp64"] + st0["tables_ms_fp32"],
                        table_bytes_fp64=st0["table_bytes_fp64"], table_bytes_fp32=st0["table_bytes_fp32"]),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
