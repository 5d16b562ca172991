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


def table_bytes_per_pair(L: int, elem: int) -> int:
    """Minimal rotation + axial table bytes of one unordered pair (equal
    orders L): sum_{n<=L}(2n+1)^2 reals + 2 sum_m (L-m1+1)^2 complex."""
    rot = sum((2 * n + 1) ** 2 for n in range(1, L + 1))
    ax = sum((L - max(1, abs(m)) + 1) ** 2 for m in range(-L, L + 1))
    return rot * elem + 2 * ax * 2 * elem


def tile_flops_per_pair(L: int, nrhs: int = 2) -> int:
    """Algorithmic flops of both directions of one pair for nrhs columns:
    16 (S_rot + S_ax) per direction and column (equal orders)."""
    s_rot = sum((2 * n + 1) ** 2 for n in range(1, L + 1))
    s_ax = sum((L - max(1, abs(m)) + 1) ** 2 for m in range(-L, L + 1))
    return 2 * nrhs * 16 * (s_rot + s_ax)


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsphere", type=int, default=5000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tight", action="store_true", help="skip the tight FP64 reference solve")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many MiB (rocprof FETCH/WRITE calibration)")
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

    from mstm_amd import geometry as G
    from mstm_amd.solver import FP32, FP64, SOLVE_LEARNED, SOLVE_MIXED, SOLVE_PLAIN, HipContext, qtotcalc

    t0 = time.time()
    cl = G.rsa_in_sphere(args.nsphere, 2.0, 0.30, seed=5, ri=1.31 + 0.0j)
    t_geom = time.time() - t0
    ctx = HipContext(0)
    order, _ = ctx.set_spheres(cl.pos, cl.radius, cl.ri, 1.0, 1e-6)
    L = int(order.max())
    ne = ctx.number_eqns
    print(f"[bench] {cl.n} spheres, order {L}, neqns {ne}", file=sys.stderr, flush=True)
    t0 = time.time()
    ctx.setup_direct(FP64)
    ctx.setup_direct(FP32)
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
    print(f"[bench] mixed (no network): {mixed_s:.3f} s, iters {rm['iterations']}", file=sys.stderr, flush=True)

    # ---- plain FP64 path (once) ----------------------------------------------
    plain = None
    if not args.no_plain:
        ctx.clear(clear_tables=False)
        t0 = time.perf_counter()
        rp = ctx.fixedorsoln(eps=1e-6, niter=10000, solver=SOLVE_PLAIN)
        plain_s = time.perf_counter() - t0
        st_plain = ctx.stats()
        q_plain = qtotcalc(cl.radius, rp["qeff"])
        plain = dict(s_per_solve=plain_s, iterations_last_pol=rp["iterations"], maxerr=rp["maxerr"],
                     qext=q_plain[0, 0], qabs=q_plain[0, 1],
                     tile_ms_avg=st_plain["tile_ms_fp64"] / max(1, st_plain["tile_launches_fp64"]),
                     iters_fp64=st_plain["iters_fp64"])
        print(f"[bench] plain FP64 solve: {plain_s:.3f} s, iters {rp['iterations']}", file=sys.stderr, flush=True)

    # ---- tight FP64 reference for the efficiency errors ----------------------
    qerr = None
    if not args.no_tight:
        ctx.clear(clear_tables=False)
        rt = ctx.fixedorsoln(eps=1e-10, niter=10000, solver=SOLVE_PLAIN)
        q_ref = qtotcalc(cl.radius, rt["qeff"])
        qerr = {}
        for name, q in (("learned", q_lig), ("mixed", q_mixed)) + ((("plain", q_plain),) if plain else ()):
            qerr[name] = dict(qext=abs(q[0, 0] - q_ref[0, 0]) / abs(q_ref[0, 0]),
                              qsca=abs(q[0, 2] - q_ref[0, 2]) / abs(q_ref[0, 2]),
                              qabs_abs=abs(q[0, 1] - q_ref[0, 1]))
        qerr["reference"] = dict(eps=1e-10, qext=q_ref[0, 0], qabs=q_ref[0, 1], qsca=q_ref[0, 2])
        print(f"[bench] tight FP64 reference done", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (FP32 tile kernel) ------------------
    npairs = cl.n * (cl.n - 1) // 2
    vec_bytes = 2 * 2 * ne * 8
    alg32 = npairs * table_bytes_per_pair(L, 4) + vec_bytes
    alg64 = npairs * table_bytes_per_pair(L, 8) + 2 * 2 * ne * 16
    t32 = st_lig["tile_ms_fp32"] / max(1, st_lig["tile_launches_fp32"])
    ach32 = alg32 / (t32 * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(HERE, "profiles", "tile_fp32_cfg5_pmc.json")
    if os.path.exists(pmc):
        traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
    roof = dict(kernel="direct_tile_kernel<float,L>", bound="hbm", achieved=round(ach32, 1), peak=HBM_PEAK_GBPS,
                unit="GB/s", frac=round(ach32 / HBM_PEAK_GBPS, 4), traffic=traffic, algorithmic_bytes=alg32,
                avg_launch_ms=round(t32, 4), flops_per_launch=npairs * tile_flops_per_pair(L))
    matvec = dict(fp32_tile_ms=round(t32, 4), fp32_hbm_GBps=round(ach32, 1))
    if plain is not None:
        t64 = plain["tile_ms_avg"]
        matvec.update(fp64_tile_ms=round(t64, 4), fp64_hbm_GBps=round(alg64 / (t64 * 1e-3) / 1e9, 1))

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
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(plain["s_per_solve"] / lig_s, 3) if plain else None,
        "learned": dict(iterations_last_pol=res["iterations"], maxerr=res["maxerr"], qext=q_lig[0, 0],
                        qabs=q_lig[0, 1], iters_fp32=st_lig["iters_fp32"] / args.steps,
                        iters_fp64_matvecs=st_lig["matvecs_fp64"] / args.steps,
                        refinements=st_lig["refinements"] / args.steps,
                        fallback_rate=st_lig["lig_fallbacks"] / lsolves,
                        start_residual_learned=st_lig["lig_err0"], start_residual_default=st_lig["lig_errb"],
                        network_ms=st_lig["lig_ms"], network=dict(samples=meta.get("samples"),
                                                                 configs=meta.get("configs"),
                                                                 loss=meta.get("loss_final"))),
        "mixed": dict(s_per_solve=round(mixed_s, 4), iterations_last_pol=rm["iterations"], maxerr=rm["maxerr"],
                      qext=q_mixed[0, 0], qabs=q_mixed[0, 1], iters_fp32=st_mixed["iters_fp32"] / len(mtimes),
                      iters_fp64_matvecs=st_mixed["matvecs_fp64"] / len(mtimes),
                      refinements=st_mixed["refinements"] / len(mtimes),
                      speedup_vs_plain_fp64=round(plain["s_per_solve"] / mixed_s, 3) if plain else None),
        "q_error_vs_tight_fp64": qerr,
        "matvec": matvec,
        "setup_s": dict(geometry=round(t_geom, 2), tables_fp64_fp32=round(t_tables, 3),
                        tables_fp64_ms=st0["tables_ms_fp64"], tables_fp32_ms=st0["tables_ms_fp32"],
                        table_bytes_fp64=st0["table_bytes_fp64"], table_bytes_fp32=st0["table_bytes_fp32"]),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
