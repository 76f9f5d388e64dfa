#!/usr/bin/env python3
"""Benchmark of the FP64 Wilson-clover dslash on one MI355X.

Workload (the drop-in GPU path configuration): SU(3) Wilson-clover, csw = 1.0,
kappa = 0.13, 16^3 x 32, time-antiperiodic, links U = exp(i eps H) at
eps = 0.5 from a seeded generator, Gaussian complex<double> spinors.  A step
is one full-lattice FP64 clover apply Q psi = gamma5 (A psi - kappa H psi)
(DiracOperator::multiply of ImprovedDiracWilsonOperator).  Inputs are resident
in HBM; ROTATE independent (links, clover, in, out) sets are cycled so the
working set (~604 MB) exceeds the 256 MB Infinity Cache.

value = algorithmic bytes / time, with 1536 B per site
  (4 links x 18 reals + spinor in 24 + out 24 + clover 72 reals, FP64).
Also reported: the even-odd Schur apply, FP64 solves per second to true
residual 1e-10 (CG on Q^2 and the even-odd BiCGStab driver), and the CPU
oracle (the FP64 restatement of the reference loops) timed on the host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-solve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "dslash GB/s (% HBM peak); CG solves/s at 1e-10, mixed vs FP64; p99 latency"
HBM_PEAK_GBS = 8000.0
DIMS = (16, 16, 16, 32)
KAPPA, CSW, EPS = 0.13, 1.0, 0.5
GAUGE_SEED, SOURCE_SEED = 1001, 2001
ROTATE = 3
BYTES_PER_SITE_FULL = (4 * 18 + 24 + 24 + 72) * 8      # 1536
BYTES_PER_ODD_SITE_SCHUR = (264 + 16 * 18) * 8         # 4416


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def fold_bc(dims, U):
    U = U.copy()
    t = np.arange(int(np.prod(dims))) % dims[3]
    U[t == dims[3] - 1, 3] *= -1
    return U


def gpu_bench(args):
    import leonardym_amd as lym

    if lym.device_count() < 1:
        raise SystemExit("no HIP device visible")
    V = int(np.prod(DIMS))
    ctx = lym.Context(DIMS, lym.REP_SU3_FUND, -1)
    sets = []
    for r in range(ROTATE):
        U = fold_bc(DIMS, lym.gen_su3_links(DIMS, EPS, GAUGE_SEED + r))
        g = lym.Gauge(ctx, U, folded=True, storage=lym.LINKS_FULL)
        op = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, CSW, True, gauge=g)
        src = lym.Spinor(ctx, lym.gaussian_spinor(V, SOURCE_SEED + r))
        out = lym.Spinor(ctx)
        eo = lym.Operator(ctx, lym.OP_EO_CLOVER, 1, KAPPA, CSW, True, gauge=g)
        sets.append((g, op, src, out, eo))
    log("bench: %s, %d rotating sets, warmup %d, steps %d" % (DIMS, ROTATE, args.warmup, args.steps))
    for i in range(args.warmup):
        g, op, src, out, eo = sets[i % ROTATE]
        op.multiply(out, src)
    ctx.synchronize()
    ctx.timer_start()
    t0 = time.perf_counter()
    for i in range(args.steps):
        g, op, src, out, eo = sets[i % ROTATE]
        op.multiply(out, src)
    ms = ctx.timer_stop()
    wall = time.perf_counter() - t0
    ms_per = ms / args.steps
    gbs = BYTES_PER_SITE_FULL * V / (ms_per * 1e-3) / 1e9
    log("bench: full clover apply %.4f ms  %.1f GB/s  (host wall %.4f ms/step)" % (ms_per, gbs, wall * 1e3 / args.steps))
    res = {"ms_per_step": ms_per, "gbs": gbs, "V": V}

    # even-odd Schur apply (two half-lattice kernels, out_e = in_e copy included)
    for i in range(args.warmup):
        g, op, src, out, eo = sets[i % ROTATE]
        eo.multiply(out, src)
    ctx.synchronize()
    ctx.timer_start()
    n_eo = max(1, args.steps // 2)
    for i in range(n_eo):
        g, op, src, out, eo = sets[i % ROTATE]
        eo.multiply(out, src)
    ms = ctx.timer_stop() / n_eo
    res["eo_schur"] = {"ms": ms, "gbs": BYTES_PER_ODD_SITE_SCHUR * (V // 2) / (ms * 1e-3) / 1e9}
    log("bench: EO Schur apply %.4f ms  %.1f GB/s" % (ms, res["eo_schur"]["gbs"]))

    if not args.no_solve:
        res["solves"] = solve_bench(lym, ctx, sets, V)
    if args.calibrate:
        # k_copy over a full 16^3x32 spinor: 24 * stride * 16 B read and written
        g, op, src, out, eo = sets[0]
        for _ in range(20):
            out.copy_from(src)
        ctx.synchronize()
        log("calibrate: 20 copies of %d bytes each" % (24 * (V // 2) * 16))
    return res


def true_residual(lym, ctx, g, x, b):
    D = lym.Operator(ctx, lym.OP_CLOVER, 1, KAPPA, CSW, False, gauge=g)
    Dx = lym.Spinor(ctx)
    D.multiply(Dx, x)
    return np.sqrt(Dx.difference_norm(b) / b.squared_norm())


def solve_bench(lym, ctx, sets, V):
    out = {}
    g, op, src, _, _ = sets[0]
    bn = src.squared_norm()
    # CG on Q^2 with rhs Q(gamma5 b): solves D x = b
    q2 = lym.Operator(ctx, lym.OP_CLOVER, 2, KAPPA, CSW, True, gauge=g)
    g5b = lym.Spinor(ctx)
    g5b.copy_from(src)
    g5b.gamma5()
    rhs = lym.Spinor(ctx)
    op.multiply(rhs, g5b)
    x = lym.Spinor(ctx)
    cg = lym.Solver(ctx, lym.SOLVER_CG, precision=(1e-11 ** 2) * rhs.squared_norm(), max_steps=10000)
    cg.solve(q2, rhs, x)  # warm
    nrep = 3
    t0 = time.perf_counter()
    for _ in range(nrep):
        info = cg.solve(q2, rhs, x)
    dt = (time.perf_counter() - t0) / nrep
    tr = true_residual(lym, ctx, g, x, src)
    out["fp64_cg_q2"] = {"solves_per_s": 1.0 / dt, "seconds": dt, "iterations": int(info.steps),
                         "true_residual": float(tr), "converged": bool(info.converged)}
    log("bench: FP64 CG on Q^2: %.3f s/solve, %d iterations, true residual %.2e" % (dt, info.steps, tr))
    # even-odd BiCGStab driver
    bi = lym.Solver(ctx, lym.SOLVER_PRECOND_BICGSTAB, precision=(1e-11 ** 2) * bn, max_steps=10000)
    bi.solve(op, src, x)
    t0 = time.perf_counter()
    for _ in range(nrep):
        info = bi.solve(op, src, x)
    dt = (time.perf_counter() - t0) / nrep
    tr = true_residual(lym, ctx, g, x, src)
    out["fp64_eo_bicgstab"] = {"solves_per_s": 1.0 / dt, "seconds": dt, "iterations": int(info.steps),
                               "true_residual": float(tr), "converged": bool(info.converged)}
    log("bench: FP64 EO BiCGStab: %.3f s/solve, %d iterations, true residual %.2e" % (dt, info.steps, tr))
    return out


def cpu_baseline(budget_s=10.0):
    """The oracle's FP64 restatement of ImprovedDiracWilsonOperator::multiply,
    timed on the host cores on the same workload (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import leonardym_amd as lym
    from helpers import K_CLOVER, OracleField, oracle

    V = int(np.prod(DIMS))
    U = fold_bc(DIMS, lym.gen_su3_links(DIMS, EPS, GAUGE_SEED))
    fld = OracleField(DIMS, U, kappa=KAPPA, csw=CSW, need_clover=True)
    psi = lym.gaussian_spinor(V, SOURCE_SEED)
    fld.apply(K_CLOVER, psi)
    t0 = time.perf_counter()
    fld.apply(K_CLOVER, psi)
    t1 = time.perf_counter() - t0
    n = int(max(3, min(300, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(n):
        fld.apply(K_CLOVER, psi)
    per = (time.perf_counter() - t0) / n
    cores = int(oracle().orc_num_threads())
    return {"value": BYTES_PER_SITE_FULL * V / per / 1e9, "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": "%d FP64 clover applies on 16^3x32 (%.1f ms each) with the oracle, OpenMP %d threads"
                      % (n, per * 1e3, cores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run 20 plain spinor copies (known bytes) to scale the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "a %d-GPU run is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    res = gpu_bench(args)
    line = {
        "metric": METRIC,
        "value": round(res["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded exp(i eps H) SU(3) links, Gaussian spinors)",
        "config": {"workload": "FP64 Wilson-clover full-lattice apply, SU(3), 16^3x32, csw=1.0, kappa=0.13, "
                               "eps=0.5, antiperiodic t, %d rotating sets" % ROTATE,
                   "lattice": "16x16x16x32", "link_storage": "18 reals", "bytes_per_site": BYTES_PER_SITE_FULL},
        "roofline": {"bound": "hbm", "achieved": round(res["gbs"], 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(res["gbs"] / HBM_PEAK_GBS, 4), "traffic": None},
        "eo_schur": {k: round(v, 4) for k, v in res["eo_schur"].items()},
    }
    if "solves" in res:
        line["solves"] = res["solves"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
