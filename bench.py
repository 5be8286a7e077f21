"""Benchmark: batched MPRGP contact-QP solves on one MI355X.

Workload (one "step"): 256 independent beam QPs (DoF log-uniform in
[5k, 60k], E log-uniform [1e4, 1e6], nu uniform [0.3, 0.48], 1-4 seeded
obstacle planes) advanced together from their warm start to KKT residual
||phi + beta|| <= 1e-6.  Matrices, right-hand sides, rows and warm starts are
assembled on the device before timing and stay resident; alpha_bar = 2/|A| is
estimated once per matrix (power method) before timing and passed in.

The headline ``value`` is the batched mixed FP32/FP64 form (``--precision``
switches it to the plain FP64 form).  Also measured on the same QPs: the
batched FP64 form, the plain FP64 form one QP at a time (p50/p99 latency,
solves/s), the mixed form's error against the FP64 results (every QP against
the GPU FP64 solve, the CPU sample against the oracle), the CPU oracle on a
bounded sample (single core), and the roofline of the solve kernel from
algorithmic bytes and device time.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_child(args):
    """Run under rocprofv3 --pmc by measure_traffic: the copy probe of known
    size, one batched solve of the benchmark workload, or the FP64 batched
    SpMV leg (power method to its cap), all with direct launches."""
    import mprgp_amd as M
    ctx = M.Context(0)
    if args.pmc_child == "copy":
        ctx.copy_probe(COPY_BYTES, 2)
        return
    from mprgp_amd import problems, workloads
    scenes = problems.batch_scenes(args.qps, seed=4, tol=args.tol, max_it=4000)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    if args.pmc_child == "spmv":
        per = [M.params(tol=1e-6, max_it=0, power_tol=1e-300, precision=M.FP64, tiles=args.ts or 4,
                        graph_steps=-16, xcd_layout=args.xcd_layout, slices_per_wave=args.slices_per_wave)
               for _ in qps]
        rc, stats = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
        print(json.dumps({"launches": stats[0]["launches"]}), flush=True)
        return
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=args.tol, max_it=0, power_tol=1e-6))
    prec = M.MIXED if args.precision == "mixed" else M.FP64
    per = [M.params(tol=args.tol, max_it=4000, alpha_bar=s["alpha_bar"], project_first=1, precision=prec,
                    tiles=args.ts or 4, graph_steps=-16, tail_tiles=args.tail_tiles, xcd_layout=args.xcd_layout,
                    slices_per_wave=args.slices_per_wave)
           for s in st]
    rc, stats = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
    print(json.dumps({"launches": stats[0]["launches"]}), flush=True)


COPY_BYTES = 1 << 28
# Read requests from the L2 to the fabric by size (32, 64, 128 bytes), then
# all of them: four TCC counters, one rocprofv3 pass.
RDREQ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")


def measure_traffic(args, tiles, what="solve"):
    """HBM bytes per launch of the step kernel at the headline's launch shape
    (``tiles`` wavefronts per workgroup) in one batched solve (``what`` =
    "solve") or in the FP64 batched SpMV leg ("spmv"): two rocprofv3 --pmc
    passes (FETCH_SIZE, then WRITE_SIZE) over a child run of the workload,
    each scaled by what the same counter reads for the copy probe
    (COPY_BYTES read and written per dispatch).  The children run while this
    process waits.  Returns None when rocprofv3 is unavailable."""
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(rocprof):
        return None
    tmp = tempfile.mkdtemp(prefix="mprgp_pmc_")
    # both unroll variants of the solve's kernel (mprgp_step<TS, MX, 3 or 6>)
    mixed = args.precision == "mixed" and what == "solve"
    kernel = f"mprgp_step<{tiles}, {'true' if mixed else 'false'},"

    def run(counter, child, tag):
        d = os.path.join(tmp, tag)
        cmd = ["timeout", "-s", "KILL", "300", rocprof, "--pmc", *counter.split(","), "--output-format", "csv",
               "-d", d, "-o", "run", "--", sys.executable, "-u", os.path.abspath(__file__),
               "--pmc-child", child, "--qps", str(args.qps), "--tol", str(args.tol),
               "--precision", args.precision, "--ts", str(tiles), "--tail-tiles", str(args.tail_tiles),
               "--xcd-layout", str(args.xcd_layout), "--slices-per-wave", str(args.slices_per_wave)]
        env = dict(os.environ, MPRGP_NO_GRAPH="1")
        r = subprocess.run(cmd, capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"rocprofv3 {counter} {child} failed ({r.returncode}): {r.stderr[-400:]}")
        launches = None
        for line in r.stdout.splitlines():
            if line.startswith("{"):
                launches = json.loads(line)["launches"]
        vals = {}
        for root, _, files in os.walk(d):
            for fn in files:
                if fn.endswith("counter_collection.csv"):
                    with open(os.path.join(root, fn)) as f:
                        for row in csv.DictReader(f):
                            name = row["Kernel_Name"]
                            if (child == "copy" and "copy_stream" in name) or (child != "copy" and kernel in name):
                                vals.setdefault(row["Counter_Name"], []).append(float(row["Counter_Value"]))
        out = {}
        for k, v in vals.items():
            if child != "copy" and launches:
                v = v[-launches:]  # the measured call's launches (the solve's power method runs first)
            out[k] = sum(v) / len(v)
        return out

    def size(counter, child, tag):  # a *_SIZE counter, KiB per dispatch -> bytes
        v = run(counter, child, tag)
        return 1024.0 * v[counter] if counter in v else None

    def by_request_size(child, tag):
        """Read bytes per dispatch from the read requests counted by size
        (32, 64 and 128 bytes), an estimate that needs no calibration."""
        v = run(",".join(RDREQ_COUNTERS), child, tag)
        if not all(c in v for c in RDREQ_COUNTERS):
            log(f"request-size counters missing for {child}: got {sorted(v)}")
            return None
        b32, b64, b128, total = (v[c] for c in RDREQ_COUNTERS)
        return {"bytes": int(32 * b32 + 64 * b64 + 128 * b128), "requests": int(total),
                "requests_by_size": [int(b32), int(b64), int(b128)]}

    try:
        cal_f = COPY_BYTES / size("FETCH_SIZE", "copy", "cf")
        cal_w = COPY_BYTES / size("WRITE_SIZE", "copy", "cw")
        fetch = size("FETCH_SIZE", what, "sf") * cal_f
        write = size("WRITE_SIZE", what, "sw") * cal_w
    except (RuntimeError, TypeError, ZeroDivisionError) as e:
        log(f"traffic measurement skipped: {e}")
        shutil.rmtree(tmp, ignore_errors=True)
        return None
    out = {"fetch_bytes": int(fetch), "write_bytes": int(write), "fetch_scale": round(cal_f, 4),
           "write_scale": round(cal_w, 4)}
    try:  # independent check of the calibrated read bytes
        copy_rq = by_request_size("copy", "cq")
        rq = by_request_size(what, "sq")
        if copy_rq and rq:
            out["fetch_by_request_size"] = {
                "bytes": rq["bytes"], "vs_calibrated_fetch": round(rq["bytes"] / fetch, 4),
                "requests_by_size_32_64_128": rq["requests_by_size"],
                "unsized_requests": rq["requests"] - sum(rq["requests_by_size"]),
                "copy_probe_bytes_vs_true": round(copy_rq["bytes"] / COPY_BYTES, 4)}
    except RuntimeError as e:
        log(f"request-size pass skipped: {e}")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def host_info():
    """The host's CPU model and core counts, reported with the CPU baseline:
    /proc/cpuinfo, os, lscpu."""
    info = {"cpu_model": None, "host_logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20)
        keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                "NUMA node(s)", "CPU max MHz")
        info["lscpu"] = {k.strip(): v.strip() for k, v in
                         (line.split(":", 1) for line in r.stdout.splitlines() if ":" in line)
                         if k.strip() in keep}
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = None
    return info


def dropin_leg(M, problems, workloads, ctx, tol=1e-8, max_it=500, calls=3):
    """The one-shot drop-in on the 133,623-DoF beam (the FP64 single-QP
    workload): what DecoupledMprgpFemSolver::forward would run each frame
    after swapping MprgpFemSolver.cpp:532-533 for mprgp_decoupled_solve --
    host CSR arrays in (A from the device assembly, downloaded once), a fresh
    context, SELL build and upload, solve, x and ev back.  The first call runs
    the power method from scratch; the later ones pass the ev of the previous
    call, as forward passes power_ev.  Beside it, the same solve with A, b,
    rows and x resident (mprgp_qp_solve)."""
    sc = problems.config2()
    dq = workloads.DeviceQP(ctx, sc)
    A = dq.fem.matrix.to_csr()
    b = dq.qp.get_rhs()
    node, normal, c = sc.rows()
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ev, ms, codes, st = np.zeros(A.shape[0]), [], [], {}
    for _ in range(calls):
        x = sc.x_n.ravel().copy()
        st = {}
        t0 = time.perf_counter()
        codes.append(M.MPRGPDecoupledCon.solve(A, b, J, c, x, tol=tol, max_it=max_it, ev=ev, stats_out=st))
        ms.append(1e3 * (time.perf_counter() - t0))
    # the drop-in's phases through the object API, same QP: a context, the
    # matrix (host SELL build + upload), the QP and its vectors, the solve
    # (alpha_bar as the drop-in's last call found it), x back
    ph = {}
    t0 = time.perf_counter()
    c2 = M.Context(0)
    ph["context"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    Am = c2.matrix_csr(A)
    ph["matrix_build_upload"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    q2 = Am.qp()
    q2.set_rows(node, normal, c)
    q2.set_rhs(b)
    q2.set_x(sc.x_n.ravel())
    ph["qp_and_vectors"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc, s2 = q2.solve(M.params(tol=tol, max_it=max_it, project_first=1, alpha_bar=st["alpha_bar"]))
    ph["solve"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    q2.get_x()
    ph["x_back"] = time.perf_counter() - t0
    c2.close()
    res = []
    for _ in range(calls):
        dq.qp.reset_x()
        t0 = time.perf_counter()
        rc, sr = dq.qp.solve(M.params(tol=tol, max_it=max_it, project_first=1, alpha_bar=st["alpha_bar"]))
        res.append(1e3 * (time.perf_counter() - t0))
    return {"workload": f"beam 100x20x20, {A.shape[0]} DoF, {A.nnz} nonzeros, {len(node)} rows, FP64, tol {tol:g}",
            "dropin_ms": [round(t, 2) for t in ms], "codes": codes, "iterations": st.get("iterations"),
            "power_iters": st.get("power_iters"),
            "phases_ms": {k: round(1e3 * v, 2) for k, v in ph.items()},
            "resident_solve_ms": [round(t, 2) for t in res], "resident_iterations": sr["iterations"],
            "note": "drop-in = host arrays in, new context, SELL build + upload, power method (no ev passed: "
                    "its 1001-iteration cap keeps ev invalid), solve, x back; resident = the same QP kept on "
                    "the device, alpha_bar given"}


def ica_leg(M, problems, workloads, ctx, tol=8e-5, max_it=5000):
    """The reference's comparison (script/data/ica_mprgp.txt: both solvers on
    the same QP at tol 8e-5): ICASolver (ICAFemSolver::forward,
    MprgpFemSolver.cpp:439-486) against the decoupled MPRGP solve, on the
    133,623-DoF beam with its plane rows, from the same start, each resident
    on the device; ICA with the reference's inner defaults (20 projected
    Gauss-Seidel sweeps to 1e-3).  MPRGP with alpha_bar from its power
    method, timed once before."""
    sc = problems.config2()
    dq = workloads.DeviceQP(ctx, sc)
    node, normal, c = sc.rows()
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    rc, st = dq.qp.solve(M.params(tol=tol, max_it=0, project_first=1))
    ab = st["alpha_bar"]
    out = {}
    for name in ("mprgp", "ica"):
        ms, its, res = [], None, None
        for _ in range(3):
            dq.qp.reset_x()
            t0 = time.perf_counter()
            if name == "mprgp":
                rc, s = dq.qp.solve(M.params(tol=tol, max_it=max_it, project_first=1, alpha_bar=ab))
            else:
                rc, s, lam, sweeps = M.ica_solve(dq.qp, J, c, tol=tol, max_it=max_it)
            ms.append(1e3 * (time.perf_counter() - t0))
            its, res = s["iterations"], s["residual"]
        out[name] = {"ms": [round(t, 2) for t in ms], "iterations": its, "residual": res, "code": rc}
    out["ica_over_mprgp_time"] = round(min(out["ica"]["ms"]) / min(out["mprgp"]["ms"]), 2)
    out["workload"] = (f"beam 100x20x20, {3 * sc.n_nodes} DoF, {len(node)} plane rows, FP64, tol {tol:g}, "
                       "both from the same start (x_n)")
    return out


def coupled_leg(M, problems, workloads, ctx, tol=1e-9, max_it=3000, reps=3):
    """The coupled-row solve (MPRGPGeneralCon<double>::solve,
    MPRGPSolver.h:594-626, with GeneralConProjector): the flat-dropped
    525-node beam with its 525 ground rows and 12 self-contact rows of 12
    nonzeros (LinearConCollider.cpp:186-217), A and b assembled on the
    device, alpha_bar from the device power method, x from the scene's
    start, projected first.  Returns the timing and, for the CPU leg, the
    same QP on the host."""
    import scipy.sparse as sp
    sc = problems.config1_flat()
    dq = workloads.DeviceQP(ctx, sc)
    node, normal, c = sc.rows()
    Jc, cc = problems.self_contact_rows(sc, np.random.default_rng(2), 12)
    J = sp.vstack([problems.rows_to_J(node, normal, sc.n_nodes), Jc]).tocsr()
    c = np.concatenate([c, cc])
    rc, st = dq.qp.solve(M.params(tol=tol, max_it=0, project_first=1))
    ab = st["alpha_bar"]
    ms, s = [], None
    for _ in range(reps):
        dq.qp.reset_x()
        t0 = time.perf_counter()
        rc, s = M.general_solve(dq.qp, J, c, tol=tol, max_it=max_it, alpha_bar=ab, project_first=1)
        ms.append(1e3 * (time.perf_counter() - t0))
    out = {"workload": f"beam 20x4x4 dropped flat, {3 * sc.n_nodes} DoF, {J.shape[0]} rows "
                       f"({len(node)} ground, {Jc.shape[0]} self-contact), FP64, tol {tol:g}",
           "gpu_ms": [round(t, 2) for t in ms], "device_ms": round(s["elapsed_ms"], 2), "code": rc,
           "iterations": s["iterations"], "steps": [s["num_cg"], s["num_exp"], s["num_prop"]]}
    host_qp = (dq.fem.matrix.to_csr(), dq.qp.get_rhs(), J, c, sc.x_n.ravel().copy(), ab, tol, max_it)
    return out, host_qp


def step_leg(M, problems, ctx, tol=1e-6, max_it=1000, newton_max_it=10, newton_eps=1e-4, reps=3):
    """One time step of MprgpFemSolver::advance (MprgpFemSolver.cpp:193-203,
    with DecoupledMprgpFemSolver::forward's Newton loop, :514-537) on the
    133,623-DoF beam, device-resident (mprgp_fem_advance): the planes'
    collision at x_n, per Newton iteration the assembly at x1 and the MPRGP
    solve, then the implicit Euler update.  alpha_bar comes from the device
    power method of an untimed first step and is passed to every timed solve;
    the CPU leg runs the oracle's step with the same alpha_bar."""
    sc = problems.config2()
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    q = fem.matrix.qp()
    planes = problems.plane_array(sc.planes)
    spar = fem.step_params(h=sc.h, gravity=sc.gravity, material=sc.material, planes=planes,
                           newton_max_it=newton_max_it, newton_eps=newton_eps)
    x_n, v_n = sc.x_n.ravel().copy(), sc.v_n.ravel().copy()
    fem.set_state(x_n, v_n)
    rc, st = fem.advance(q, spar, tol=tol, max_it=max_it)
    ab = st["last"]["alpha_bar"]
    ms = []
    for _ in range(reps):
        fem.set_state(x_n, v_n)
        t0 = time.perf_counter()
        rc, st = fem.advance(q, spar, tol=tol, max_it=max_it, alpha_bar=ab)
        ms.append(1e3 * (time.perf_counter() - t0))
    x_gpu, _ = fem.get_state()
    out = {"workload": f"beam 100x20x20, {3 * sc.n_nodes} DoF, planes collided at x_n, Newton to "
                       f"{newton_eps:g} (at most {newton_max_it}), FP64 solves to {tol:g}",
           "gpu_ms": [round(t, 2) for t in ms], "code": rc, "rows": st["rows"],
           "newton_iters": st["newton_iters"], "solve_iterations": st["iterations"]}
    host = (sc, planes, x_n, v_n, ab, tol, max_it, newton_max_it, newton_eps, x_gpu)
    return out, host


def spmv_leg(M, workloads, ctx, qps, tiles, graph_steps):
    """Bandwidth of the batched SpMV (FixedSparseMatrix::multiply,
    MPRGPSolver.h:469-473) on the whole batch: every QP runs the power method
    (specRad, MPRGPSolver.h:42-79) to its 1001-iteration cap (a power_tol of
    1e-300 is never reached), so each of the ~1000 launches multiplies every
    QP's matrix once and normalises (op POW_IT: the matrix, w read, v read and
    written, w written).  Achieved = algorithmic bytes of the POW_IT launches
    (from the op counts) / device time of the whole launch sequence (HIP events
    on the solve stream), so the few set-up launches count against it.  The
    batch's FP64 matrices total about 1.8 GB, far above the 256 MB last-level
    cache; the FP32 copies (mixed form) about 1 GB."""
    pow_it = workloads.OPS.index("POW_IT")
    out = {}
    for name, prec in (("fp64_matrix", M.FP64), ("fp32_matrix", M.MIXED)):
        mixed = prec == M.MIXED
        per = [M.params(tol=1e-6, max_it=0, power_tol=1e-300, precision=prec, tiles=tiles,
                        graph_steps=graph_steps) for _ in qps]
        _, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
        alg, its = 0, 0
        idx = max(workloads.index_bytes(q.n_nodes) for q in qps)  # the launch's column index width
        for q in qps:
            k = int(q.qp.op_counts()[pow_it])
            its = max(its, k)
            alg += k * workloads.op_bytes(pow_it, q.n_nodes, q.n_blocks, q.m, mixed=mixed, idx=idx)
        dev_s = st[0]["elapsed_ms"] / 1e3
        launches = st[0]["launches"]
        out[name] = {"achieved": round(alg / dev_s / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg / dev_s / 1e9 / HBM_PEAK_GBS, 4), "launches": launches,
                     "power_iterations": its, "avg_launch_us": round(1e6 * dev_s / max(launches, 1), 2),
                     "bytes_per_launch": int(alg / max(its, 1))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qps", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=3.0)
    ap.add_argument("--ts", type=int, default=0, help="wavefronts per workgroup (0: library choice)")
    ap.add_argument("--graph-steps", type=int, default=0)
    ap.add_argument("--tail-tiles", type=int, default=0, help="tail switch point (0: library default)")
    ap.add_argument("--xcd-layout", type=int, default=0, help="0 runs of tiles over the XCDs, 1 per-QP chunks")
    ap.add_argument("--slices-per-wave", type=int, default=1, help="SELL slices each wavefront streams per launch")
    ap.add_argument("--no-tune", action="store_true", help="library default shape instead of the autotuner's")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM counter passes")
    ap.add_argument("--no-spmv", action="store_true", help="skip the batched SpMV bandwidth leg")
    ap.add_argument("--no-dropin", action="store_true", help="skip the one-shot drop-in leg")
    ap.add_argument("--no-ica", action="store_true", help="skip the ICA-versus-MPRGP comparison leg")
    ap.add_argument("--no-coupled", action="store_true", help="skip the coupled-row solve leg")
    ap.add_argument("--no-step", action="store_true", help="skip the time-step leg")
    ap.add_argument("--pmc-child", choices=["copy", "solve", "spmv"], help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: the batched QPs are independent "
                          "and would shard across GPUs with no communication, one process per "
                          "GPU, but only --gpus 1 is measured"}))
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

    copy_gbps = ctx.copy_probe(1 << 30, 10)
    log(f"copy probe: {copy_gbps:.0f} GB/s")
    prec = M.MIXED if args.precision == "mixed" else M.FP64
    mixed = prec == M.MIXED

    def batch_params(precision):
        return [M.params(tol=args.tol, max_it=4000, alpha_bar=q.alpha_bar, project_first=1,
                         precision=precision, tiles=args.ts, graph_steps=args.graph_steps,
                         tail_tiles=args.tail_tiles, xcd_layout=args.xcd_layout,
                         slices_per_wave=args.slices_per_wave, from_x0=1) for q in qps]

    per = batch_params(prec)
    shape = None
    if not args.no_tune and not (args.ts or args.graph_steps):
        from mprgp_amd import tune
        t1 = time.perf_counter()
        per, shape, hit = tune.tuned_params(ctx, [q.qp for q in qps], per, prec, log=log)
        log(f"launch shape {shape} ({'cached' if hit else 'tuned'}, {time.perf_counter() - t1:.1f}s)")

    def batch_step(params):
        # every QP restarts from its stored warm start inside the solve
        # (from_x0: the x = x0 copies ride on the batch's set-up launch)
        return M.solve_batch(ctx, [q.qp for q in qps], per_qp=params)

    for _ in range(args.warmup):
        batch_step(per)
    times, stats = [], None
    tt0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        rc, stats = batch_step(per)
        times.append(time.perf_counter() - s0)
    total = time.perf_counter() - tt0
    conv = sum(1 for s in stats if s["code"] == 0)
    finished64 = sum(s["fp64_finish"] for s in stats)
    value = args.qps * args.steps / total
    log(f"batched {args.precision}: {value:.1f} solves/s, {1e3 * total / args.steps:.2f} ms/step, "
        f"converged {conv}/{args.qps}")
    x_head = [q.qp.get_x() for q in qps]

    tiles = shape["tiles"] if shape else (args.ts or 4)
    if shape:  # the counter passes run the tuned shape
        args.tail_tiles = int(shape.get("tail_tiles", 0))
        args.xcd_layout = int(shape.get("xcd_layout", 0))
        args.slices_per_wave = int(shape.get("slices_per_wave", 1))
    traffic = None if args.no_traffic else measure_traffic(args, tiles)
    if traffic:
        log(f"traffic per launch (mprgp_step<{tiles}, ...>): {traffic}")

    # roofline of mprgp_step over the last batched solve: algorithmic bytes /
    # device time (HIP events on the solve stream around the launches)
    launches = stats[0]["launches"]
    dev_s = stats[0]["elapsed_ms"] / 1e3
    idx = max(workloads.index_bytes(q.n_nodes, args.slices_per_wave) for q in qps)
    alg = sum(workloads.solve_bytes(q.qp.op_counts(), q.n_nodes, q.n_blocks, q.m, mixed, idx=idx) for q in qps)
    achieved = alg / dev_s / 1e9
    roof = {"kernel": "mprgp_step", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (traffic["fetch_bytes"] + traffic["write_bytes"]) if traffic else None,
            "traffic_detail": traffic, "bytes_per_launch": int(alg / max(launches, 1)), "index_bytes": idx,
            "avg_launch_us": round(1e6 * dev_s / max(launches, 1), 2), "launches": launches,
            "copy_gbps": round(copy_gbps, 1),
            "iterations_total": int(sum(s["iterations"] for s in stats))}
    log(f"roofline: {roof}")

    spmv = None
    if not args.no_spmv:
        spmv = spmv_leg(M, workloads, ctx, qps, tiles, shape["graph_steps"] if shape else args.graph_steps)
        st = None if args.no_traffic else measure_traffic(args, tiles, "spmv")
        if st:
            f64 = spmv["fp64_matrix"]
            f64["traffic"] = st["fetch_bytes"] + st["write_bytes"]
            f64["traffic_detail"] = st
            f64["traffic_vs_algorithmic"] = round(f64["traffic"] / f64["bytes_per_launch"], 3)
        log(f"batched spmv (power method to its cap): {spmv}")

    other, st64 = None, None
    if mixed and not args.no_fp64_batch:
        per64 = batch_params(M.FP64)
        if shape is not None:
            from mprgp_amd import tune
            per64, _, _ = tune.tuned_params(ctx, [q.qp for q in qps], per64, M.FP64, log=log)
        batch_step(per64)
        k = max(1, args.steps // 2)
        s0 = time.perf_counter()
        for _ in range(k):
            rc, st64 = batch_step(per64)
        t64 = time.perf_counter() - s0
        other = {"solves_per_s": round(args.qps * k / t64, 2), "ms_per_step": round(1e3 * t64 / k, 3),
                 "converged": sum(1 for s in st64 if s["code"] == 0),
                 "iterations_total": int(sum(s["iterations"] for s in st64))}
        log(f"batched fp64: {other}")

    plain, x_plain = None, None
    if not args.no_plain:
        lat, x_plain = [], []
        for q in qps:
            q.qp.reset_x()
            pp = M.params(tol=args.tol, max_it=4000, alpha_bar=q.alpha_bar, project_first=1)
            s0 = time.perf_counter()
            q.qp.solve(pp)
            lat.append(time.perf_counter() - s0)
            x_plain.append(q.qp.get_x())
        lat = np.array(lat) * 1e3
        plain = {"solves_per_s": round(len(lat) / (lat.sum() / 1e3), 2),
                 "p50_ms": round(float(np.percentile(lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat, 99)), 3)}
        log(f"plain fp64 one at a time: {plain}")

    err = None
    if mixed and x_plain is not None:
        rel = np.array([np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300) for a, b in zip(x_head, x_plain)])
        err = {"vs_gpu_fp64_max": float(rel.max()), "vs_gpu_fp64_median": float(np.median(rel))}

    step, step_host = None, None
    if not args.no_step:
        step, step_host = step_leg(M, problems, ctx)
        log(f"time step: {step}")

    coupled, coupled_host = None, None
    if not args.no_coupled:
        coupled, coupled_host = coupled_leg(M, problems, workloads, ctx)
        log(f"coupled rows: {coupled}")

    cpu, cpu_all = None, None
    if not args.no_cpu:
        import oracle
        order = np.argsort([q.n_nodes for q in qps])
        done, spent, dof, rel_o, sample, per_ms, ostats, work = 0, 0.0, 0, [], [], [], [], []
        for i in order:
            sc = scenes[i]
            tets, d, vol, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
            A, b = oracle.fem_build_system(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson,
                                           sc.density, sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity,
                                           sc.h, sc.material)
            node, normal, c = sc.rows()
            J = problems.rows_to_J(node, normal, sc.n_nodes)
            s0 = time.perf_counter()
            _, xo, so, _ = oracle.decoupled_solve(A, b, J, c, sc.x_n.ravel(), tol=args.tol, max_it=4000,
                                                  alpha_bar=qps[i].alpha_bar, project_first=True)
            dt = time.perf_counter() - s0
            spent += dt
            per_ms.append(1e3 * dt)
            ostats.append(so)
            work.append(A.nnz * (so["iterations"] + 1))
            sample.append((A, b, J, c, sc.x_n.ravel(), qps[i].alpha_bar))
            rel_o.append(np.linalg.norm(x_head[i] - xo) / np.linalg.norm(xo))
            done += 1
            dof += 3 * sc.n_nodes
            if spent > args.cpu_seconds:
                break
        hinfo = host_info()
        # the whole workload by extrapolation: seconds per (nonzero x
        # iteration) from the sample, times each QP's nonzeros and the
        # iterations of its FP64 GPU solve (the same algorithm); checked
        # against one timed solve of the largest QP
        rate = spent / float(sum(work))
        it64 = [s["iterations"] for s in st64] if other is not None else [s["iterations"] for s in stats]
        sampled = set(int(i) for i in order[:done])
        est = spent + sum(rate * qps[j].n_blocks * 9 * (it64[j] + 1) for j in range(len(qps)) if j not in sampled)
        # the QP with the most predicted work, timed to check the model
        big = max(range(len(qps)), key=lambda j: qps[j].n_blocks * (it64[j] + 1))
        scb = scenes[big]
        tets, d, vol, mass = oracle.fem_setup(scb.n_nodes, scb.tets, scb.X0)
        Ab, bb = oracle.fem_build_system(scb.n_nodes, tets, d, vol, mass, scb.young, scb.poisson, scb.density,
                                         scb.x_n.ravel(), scb.v_n.ravel(), scb.gravity, scb.h, scb.material)
        nb_, nrm_, cb_ = scb.rows()
        s0 = time.perf_counter()
        _, _, sob, _ = oracle.decoupled_solve(Ab, bb, problems.rows_to_J(nb_, nrm_, scb.n_nodes), cb_,
                                              scb.x_n.ravel(), tol=args.tol, max_it=4000,
                                              alpha_bar=qps[big].alpha_bar, project_first=True)
        big_s = time.perf_counter() - s0
        big_pred = rate * Ab.nnz * (sob["iterations"] + 1)
        cpu = {"value": round(done / spent, 3), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"{done} smallest of the {args.qps} QPs ({dof} DoF total), oracle FP64, "
                         f"same alpha_bar, {spent:.1f}s, 1 thread on {hinfo['cpu_model']}",
               "host": hinfo,
               "full_workload_estimate": {
                   "solves_per_s": round(args.qps / est, 3), "seconds": round(est, 1),
                   "model": "sample + (s per nonzero-iteration of the sample) x nonzeros x (FP64 GPU "
                            "iterations + 1) for the QPs not timed",
                   "costliest_qp": {"dof": 3 * scb.n_nodes, "measured_s": round(big_s, 2),
                                    "model_s": round(big_pred, 2), "iterations": sob["iterations"]}},
               "per_solve_ms": {"p50": round(float(np.percentile(per_ms, 50)), 3),
                                "p99": round(float(np.percentile(per_ms, 99)), 3)},
               "steps": {k: int(sum(o[k] for o in ostats))
                         for k in ("iterations", "num_cg", "num_exp", "num_prop")}}
        if err is None:
            err = {}
        err["vs_oracle_max"] = float(max(rel_o))
        err["vs_oracle_sample"] = done
        log(f"cpu oracle: {cpu}")
        # the same sample over all host cores of this GPU's share: independent
        # QPs solved concurrently (the oracle's ctypes calls release the GIL)
        import concurrent.futures as cf
        # every CPU this GPU's share of the host gives the process
        # (OMP_NUM_THREADS caps it to that share where it is set)
        usable = hinfo["usable_cpus"] or os.cpu_count() or 1
        workers = max(1, min(usable, int(os.environ.get("OMP_NUM_THREADS") or usable)))

        def one(item):
            A, b, J, c, x0, ab = item
            return oracle.decoupled_solve(A, b, J, c, x0, tol=args.tol, max_it=4000, alpha_bar=ab,
                                          project_first=True)[0]

        s0 = time.perf_counter()
        with cf.ThreadPoolExecutor(workers) as pool:
            list(pool.map(one, sample))
        wall = time.perf_counter() - s0
        cpu_all = {"value": round(done / wall, 3), "unit": "solves/s", "cores": workers, "kind": "port",
                   "sample": f"the same {done} QPs, one per thread at a time, {wall:.1f}s wall, "
                             f"{workers} threads = the CPUs this process may use (host: "
                             f"{hinfo['host_logical_cpus']} logical CPUs)"}
        del sample
        log(f"cpu oracle, all cores: {cpu_all}")
        if coupled_host is not None:
            # the coupled solve's restatement on the same QP, one thread
            A_, b_, J_, c_, x0_, ab_, tol_, it_ = coupled_host
            s0 = time.perf_counter()
            code_, _, so_ = oracle.general_solve(A_, b_, J_, c_, x0_, tol=tol_, max_it=it_, alpha_bar=ab_)
            coupled["cpu_oracle"] = {"ms": round(1e3 * (time.perf_counter() - s0), 1), "code": code_,
                                     "iterations": so_["iterations"], "cores": 1, "kind": "port"}
            log(f"coupled rows, cpu oracle: {coupled['cpu_oracle']}")
        if step_host is not None:
            # the oracle's time step (collision, Newton loop with assembly and
            # solve, implicit Euler) from the same state, one thread
            sc_, planes_, x_n_, v_n_, ab_, tol_, it_, nmax_, neps_, x_gpu_ = step_host
            tets_, d_, vol_, mass_ = oracle.fem_setup(sc_.n_nodes, sc_.tets, sc_.X0)
            s0 = time.perf_counter()
            x_o, _, info_ = oracle.advance(sc_.n_nodes, tets_, d_, vol_, mass_, sc_.young, sc_.poisson,
                                           sc_.density, x_n_, v_n_, sc_.gravity, sc_.h, sc_.material, planes_,
                                           newton_max_it=nmax_, newton_eps=neps_, tol=tol_, max_it=it_,
                                           alpha_bar=ab_)
            step["cpu_oracle"] = {"ms": round(1e3 * (time.perf_counter() - s0), 1), "rows": info_["rows"],
                                  "newton_iters": info_["newton_iters"], "solve_iterations": info_["iterations"],
                                  "cores": 1, "kind": "port",
                                  "x_rel_diff_vs_gpu": float(np.linalg.norm(x_gpu_ - x_o) / np.linalg.norm(x_o))}
            log(f"time step, cpu oracle: {step['cpu_oracle']}")

    serve = None
    if not args.no_serving:
        from mprgp_amd import serving
        ss = serving.make_sessions(ctx, count=64, seed=5, own_contexts=True)
        eng = M.Engine(ctx, max_active=256, tiles=4, hbm_budget=int(200e9))
        serving.run_open_loop(eng, ss, rate=2000.0, duration=0.3, precision=prec, seed=105)  # warm-up
        serve = serving.run_open_loop(eng, ss, rate=2000.0, duration=args.serving_seconds,
                                      precision=prec, seed=6)
        serve.pop("engine", None)
        serve["sessions"] = len(ss)
        serve["dof_total"] = int(sum(3 * s.n_nodes for s in ss))
        serve["fp64_error_sample"] = serving.fp64_error_sample(eng, ss, n=16, precision=prec)
        serve["plain_fp64_one_at_a_time"] = serving.run_one_at_a_time(ss, n_req=150, seed=7)
        eng.close()
        log(f"serving: {serve}")

    dropin = None
    if not args.no_dropin:
        dropin = dropin_leg(M, problems, workloads, ctx)
        log(f"drop-in: {dropin}")
    ica = None
    if not args.no_ica:
        ica = ica_leg(M, problems, workloads, ctx)
        log(f"ica vs mprgp: {ica}")

    out = {
        "metric": "MPRGP solves/s at KKT res \u22641e-6, batched FP32/FP64 vs plain FP64; p50/p99 ms",
        "value": round(value, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not mixed else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "256 beam contact QPs batched, DoF 5k-60k each", "qps": args.qps,
                   "launch_shape": shape,
                   "dof_total": 3 * total_nodes, "blocks_total": total_blocks,
                   "dof_range": [5000, 60000], "tol": args.tol, "precision": args.precision,
                   "converged": conv, "fp64_finish": finished64},
        "batch_p50_ms": round(1e3 * float(np.percentile(times, 50)), 3),
        "batch_p99_ms": round(1e3 * float(np.percentile(times, 99)), 3),
        "batched_fp64": other,
        "plain_fp64_one_at_a_time": plain,
        "speedup_vs_plain": round(value / plain["solves_per_s"], 2) if plain else None,
        "solution_rel_err": err,
        "serving": serve,
        "roofline": roof,
        "spmv_roofline": spmv,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "dropin_fp64_133k": dropin,
        "ica_vs_mprgp_133k": ica,
        "coupled_rows_537": coupled,
        "time_step_133k": step,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
