"""Benchmark: batched MPRGP contact-QP solves on one MI355X.

Workload (one "step"): 256 independent beam QPs (DoF log-uniform in
[5k, 60k], E log-uniform [1e4, 1e6], nu uniform [0.3, 0.48], 1-4 seeded
obstacle planes) advanced together from their warm start to KKT residual
||phi + beta|| <= 1e-6.  Matrices, right-hand sides, rows and warm starts are
assembled on the device before timing and stay resident; alpha_bar = 2/|A| is
estimated once per matrix (power method) before timing and passed in.

Also measured: the plain FP64 form solving the same QPs one at a time
(p50/p99 latency, solves/s), the CPU oracle on a bounded sample of them
(single core), and the roofline of the solve kernel from algorithmic bytes.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qps", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ts", type=int, default=0, help="wavefronts per workgroup (0: library choice)")
    ap.add_argument("--spw", type=int, default=0, help="slices per wavefront per tile (0: library choice)")
    ap.add_argument("--graph-steps", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the batched QPs shard across "
                          "GPUs with no communication (see DESIGN.md), but only --gpus 1 is run"}))
        sys.exit(2)

    import mprgp_amd as M
    from mprgp_amd import problems, workloads

    t0 = time.perf_counter()
    scenes = problems.batch_scenes(args.qps, seed=4, tol=args.tol, max_it=4000)
    ctx = M.Context(0)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    total_nodes = sum(q.n_nodes for q in qps)
    total_blocks = sum(q.n_blocks for q in qps)
    log(f"setup: {len(qps)} QPs, {3 * total_nodes} DoF, {total_blocks} blocks, "
        f"{time.perf_counter() - t0:.1f}s")

    # alpha_bar per matrix: power method (reference specRad), batched, before timing
    t1 = time.perf_counter()
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=args.tol, max_it=0, power_tol=1e-6))
    for q, s in zip(qps, st):
        q.alpha_bar = s["alpha_bar"]
    log(f"power method: {time.perf_counter() - t1:.2f}s, "
        f"iterations max {max(s['power_iters'] for s in st)}")

    prec = M.MIXED if args.precision == "mixed" else M.FP64
    per = [M.params(tol=args.tol, max_it=4000, alpha_bar=q.alpha_bar, project_first=1,
                    precision=prec, tiles=args.ts, graph_steps=args.graph_steps) for q in qps]
    for p in per:
        p.reserved_i[0] = args.spw

    def batch_step():
        for q in qps:
            q.qp.reset_x()
        return M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)

    for _ in range(args.warmup):
        batch_step()
    times, stats = [], None
    tt0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        rc, stats = batch_step()
        times.append(time.perf_counter() - s0)
    total = time.perf_counter() - tt0
    conv = sum(1 for s in stats if s["code"] == 0)
    dev_ms = [s["elapsed_ms"] for s in stats][0]
    value = args.qps * args.steps / total
    log(f"batched: {value:.1f} solves/s, {1e3 * total / args.steps:.2f} ms/step, converged {conv}/{args.qps}")

    # roofline of mprgp_step over one batched solve: algorithmic bytes / device time
    launches = stats[0]["launches"]
    alg = 0
    s_bytes = 8
    for q, s in zip(qps, stats):
        cnt = q.qp.op_counts()
        alg += workloads.solve_bytes(cnt, q.n_nodes, q.n_blocks, q.m, s_bytes)
    dev_s = dev_ms / 1e3
    achieved = alg / dev_s / 1e9
    roof = {"kernel": "mprgp_step", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "bytes_per_launch": int(alg / max(launches, 1)),
            "avg_launch_us": round(1e6 * dev_s / max(launches, 1), 2), "launches": launches}

    plain = None
    if not args.no_plain:
        lat = []
        for q, p in zip(qps, per):
            q.qp.reset_x()
            pp = M.params(tol=p.tol, max_it=p.max_it, alpha_bar=p.alpha_bar, project_first=1)
            s0 = time.perf_counter()
            q.qp.solve(pp)
            lat.append(time.perf_counter() - s0)
        lat = np.array(lat) * 1e3
        plain = {"solves_per_s": round(len(lat) / (lat.sum() / 1e3), 2),
                 "p50_ms": round(float(np.percentile(lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat, 99)), 3)}
        log(f"plain fp64 one at a time: {plain}")

    cpu = None
    if not args.no_cpu:
        import oracle
        order = np.argsort([q.n_nodes for q in qps])
        done, spent, dof = 0, 0.0, 0
        for i in order:
            sc = scenes[i]
            tets, d, vol, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
            A, b = oracle.fem_build_system(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson,
                                           sc.density, sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity,
                                           sc.h, sc.material)
            node, normal, c = sc.rows()
            J = problems.rows_to_J(node, normal, sc.n_nodes)
            s0 = time.perf_counter()
            oracle.decoupled_solve(A, b, J, c, sc.x_n.ravel(), tol=args.tol, max_it=4000,
                                   alpha_bar=qps[i].alpha_bar, project_first=True)
            spent += time.perf_counter() - s0
            done += 1
            dof += 3 * sc.n_nodes
            if spent > args.cpu_seconds:
                break
        cpu = {"value": round(done / spent, 3), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"{done} smallest of the {args.qps} QPs ({dof} DoF total), oracle FP64, "
                         f"same alpha_bar, {spent:.1f}s; host {platform.processor() or platform.machine()}"}
        log(f"cpu oracle: {cpu}")

    out = {
        "metric": "MPRGP solves/s at KKT res <=1e-6 (batched, 256 QPs)",
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == M.FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "batched beams (configs[3])", "qps": args.qps,
                   "dof_total": 3 * total_nodes, "blocks_total": total_blocks,
                   "dof_range": [5000, 60000], "tol": args.tol, "precision": args.precision,
                   "converged": conv},
        "batch_p50_ms": round(1e3 * float(np.percentile(times, 50)), 3),
        "batch_p99_ms": round(1e3 * float(np.percentile(times, 99)), 3),
        "plain_fp64_one_at_a_time": plain,
        "speedup_vs_plain": round(value / plain["solves_per_s"], 2) if plain else None,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
