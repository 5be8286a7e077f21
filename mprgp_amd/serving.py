"""Serving loop: simulation sessions issuing one contact QP per timestep to
the continuous-batching engine.

A session is one deformable body (a beam or a blob, tet mesh) on seeded
obstacle planes.  Its matrix, right-hand side, rows and warm start live in one
device-resident QP.  A timestep (one request) is

* every ``reassemble`` steps: a random rotation of the body about its
  centroid (then lifted clear of the planes), full corotational assembly of
  A and b at x_n, and a fresh alpha_bar;
* otherwise: right-hand side only (``RHS_ONLY``: the lagged matrix is kept);
* every step: the collider's rows from x_n (a node within the planes' band
  gets one row, as the reference's collision handling does per step);
* the solve from the warm start x_n (projected first, as
  DecoupledMprgpFemSolver::forward does, MprgpFemSolver.cpp:528-533);
* x_{n+1} = x and updateMesh's implicit-Euler velocity,
  v_{n+1} = v_n + ((x - (x_n + h v_n)) / (h h)) h.

By default a session's positions and velocities live on the device
(``Fem.begin_step`` before the solve: collision, assembly, warm start;
``Fem.end_step`` after it: positions and velocities), and the host only
reads them back for the rotation of a re-assembly step.  With
``device_state=False`` the host keeps them, assembles from host arrays and
uploads the host collider's rows; the two kinds step bit-identically.

``run_open_loop`` drives the sessions with a Poisson stream of requests
through an :class:`mprgp_amd.Engine`; ``run_one_at_a_time`` serves the same
kind of requests with the plain FP64 solve, one QP at a time.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import dataclasses
import math
import threading
import time

import numpy as np

import mprgp_amd as M
from mprgp_amd import problems


@dataclasses.dataclass
class Session:
    sid: int
    scene: problems.Scene
    fem: M.Fem
    qp: M.QP
    alpha_bar: float = 0.0
    step: int = 0
    reassemble: int = 10
    rng: np.random.Generator = None
    history: list = dataclasses.field(default_factory=list)
    pending: collections.deque = dataclasses.field(default_factory=collections.deque)
    busy: bool = False
    device_state: bool = True
    sp: object = None  # StepParams of the open step (device state)

    @property
    def n_nodes(self):
        return self.scene.n_nodes

    def prepare(self):
        """Host + device work before the solve of the current step."""
        sc = self.scene
        full = self.step % self.reassemble == 0
        if self.device_state and self.step > 0 and full:
            x, v = self.fem.get_state()
            sc.x_n, sc.v_n = x.reshape(-1, 3), v.reshape(-1, 3)
        if full and self.step > 0:
            ang = math.radians(self.rng.uniform(1.0, 5.0))
            axis = self.rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
            R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
            c = sc.x_n.mean(axis=0)
            x = (sc.x_n - c) @ R.T + c
            # lift along +y until no node is inside a plane (normals have n_y > 0)
            lift = 0.0
            for n, p, _ in sc.planes:
                n = np.asarray(n) / np.linalg.norm(n)
                lift = max(lift, float(np.max(-(x @ n + p))) / n[1])
            x[:, 1] += max(lift, 0.0)
            sc.x_n = x
            sc.v_n = sc.v_n @ R.T
        material = M.COROTATIONAL if full else (M.COROTATIONAL | M.RHS_ONLY)
        if self.device_state:
            if self.step == 0 or full:
                self.fem.set_state(sc.x_n.ravel(), sc.v_n.ravel())
            self.sp = self.fem.step_params(h=sc.h, gravity=sc.gravity, material=material,
                                           planes=problems.plane_array(sc.planes), newton_max_it=1)
            self.fem.begin_step(self.qp, self.sp)
            return
        self.fem.assemble(self.qp, sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, material)
        if self.step == 0 or full:  # otherwise the QP's x is the last solution = x_n
            self.qp.set_x(sc.x_n.ravel())
        node, normal, c = sc.rows()
        self.qp.set_rows(node, normal, c)

    def complete(self, keep_history=False, rc=0, stats=None):
        sc = self.scene
        if stats is not None and self.step % self.reassemble == 0:
            self.alpha_bar = stats["alpha_bar"]
        if self.device_state:
            if stats is not None and not math.isfinite(stats["residual"]):
                raise FloatingPointError(f"session {self.sid} ({sc.name}, {3 * self.n_nodes} DoF) step "
                                         f"{self.step}: non-finite residual, code {rc}, stats {stats}")
            self.fem.end_step(self.qp, self.sp)
            if keep_history:
                self.history.append(self.fem.get_state()[0].reshape(-1, 3))
            self.step += 1
            return
        x = self.qp.get_x().reshape(-1, 3)
        if not np.all(np.isfinite(x)):
            raise FloatingPointError(f"session {self.sid} ({sc.name}, {3 * self.n_nodes} DoF) step "
                                     f"{self.step}: non-finite solution, code {rc}, stats {stats}")
        if keep_history:
            self.history.append(x.copy())
        # updateMesh under implicit Euler, in the reference's arithmetic
        psi = sc.x_n + sc.h * sc.v_n
        acc = (x - psi) / ((1.0 * sc.h) * sc.h)
        sc.v_n = sc.v_n + (acc * 1.0) * sc.h
        sc.x_n = x
        self.step += 1


def session_scene(rng, sid, dof_range=(3e3, 1.5e5)):
    """Seeded beam or blob with DoF log-uniform in dof_range, E log-uniform
    [1e4, 1e6], nu uniform [0.3, 0.48], falling at ~1 m/s onto 1-3 planes."""
    dof = math.exp(rng.uniform(math.log(dof_range[0]), math.log(dof_range[1])))
    blob = rng.uniform() < 0.5
    if blob:
        k = max(2, int(round((dof / 3.0) ** (1.0 / 3.0))) - 1)
        nx, ny, nz = k, k, k
    else:
        nx, ny, nz = problems.dims_for_dof(dof)
    E = math.exp(rng.uniform(math.log(1e4), math.log(1e6)))
    nu = rng.uniform(0.3, 0.48)
    sc = problems.beam_scene(f"{'blob' if blob else 'beam'}{sid}", nx, ny, nz, young=E, poisson=nu,
                             angle=rng.uniform(0, 10), material=M.COROTATIONAL, rng=rng)
    if blob:  # smooth random bulges of the rest shape
        X = sc.X0
        c = X.mean(axis=0)
        d = X - c
        r = np.linalg.norm(d, axis=1) + 1e-12
        f = 1.0 + 0.15 * np.sin(3.0 * d[:, 0] / (r.max())) * np.cos(2.0 * d[:, 2] / r.max())
        X = c + d * f[:, None]
        X[:, 1] += 0.01 - X[:, 1].min()
        sc.X0 = X
        sc.tets = problems.orient_positive(X, sc.tets)
        sc.x_n = X.copy()
    # the ground and 0-2 planes tilted up to 30 degrees, each a few mm below
    # the body: it falls onto them (no initial penetration)
    sc.planes = [((0.0, 1.0, 0.0), 0.0, 0.01)]
    for _ in range(int(rng.integers(0, 3))):
        tilt = math.radians(rng.uniform(0.0, 30.0))
        az = rng.uniform(0.0, 2 * math.pi)
        n = np.array([math.sin(tilt) * math.cos(az), math.cos(tilt), math.sin(tilt) * math.sin(az)])
        gap = rng.uniform(0.002, 0.02)
        sc.planes.append((tuple(n), gap - float(np.min(sc.x_n @ n)), 0.01))
    return sc


def make_sessions(ctx, count=64, seed=5, dof_range=(3e3, 1.5e5), reassemble=10, own_contexts=False,
                  device_state=True, only=None):
    """Sessions on `ctx`, or each on a context of its own (same device) so
    that the host work of different sessions does not share one stream.
    `only`: the session ids to build (the others' scenes are drawn and
    dropped, so session i is the same whichever others are built)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        sc = session_scene(rng, i, dof_range)
        if only is not None and i not in only:
            continue
        sctx = M.Context(0) if own_contexts else ctx
        if own_contexts:
            ctx._objs.append(sctx)
        fem = M.Fem(sctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
        qp = fem.matrix.qp()
        out.append(Session(i, sc, fem, qp, reassemble=reassemble,
                           rng=np.random.default_rng([seed, i]), device_state=device_state))
    # alpha_bar per session from its first matrix (power method, batched)
    for s in out:
        s.prepare()
    for s in out:
        rc, d = s.qp.solve(M.params(tol=1e-6, max_it=0, power_tol=1e-6))
        s.alpha_bar = d["alpha_bar"]
    return out


def solve_params(s: Session, tol, precision, tiles=4, max_it=1000):
    """The simulator's call: warm start projected first and the simulator's
    iteration cap (MprgpFemSolver.cpp:189).  A step that re-assembled the
    matrix re-estimates alpha_bar = 2/|A| by the power method started from
    the eigenvector kept from the previous estimate (the reference's
    power_ev, MprgpFemSolver.cpp:533); right-hand-side-only steps keep the
    matrix and so its alpha_bar."""
    full = s.step % s.reassemble == 0
    return M.params(tol=tol, max_it=max_it, alpha_bar=0.0 if full else s.alpha_bar, project_first=1,
                    precision=precision, tiles=tiles, use_ev=1, power_tol=1e-6)


def run_open_loop(engine: M.Engine, sessions, rate=2000.0, duration=3.0, tol=1e-6,
                  precision=M.MIXED, seed=6, workers=4, keep_history=False, max_it=1000):
    """Poisson arrivals at `rate` requests/s for `duration` s, each to a
    uniformly chosen session; a session's requests are served in order.
    Latency = completion of the solve - arrival of the request.
    keep_history: True for every session, or a set of session ids whose
    positions after every step are kept.  The result reports each solve's
    error bound (mprgp_stats.error_bound: |x - x*| <= residual / lambda_min
    for the session's mass-dominated matrix)."""
    rng = np.random.default_rng(seed)
    n_req = max(1, int(rng.poisson(rate * duration)))
    arr = np.sort(rng.uniform(0.0, duration, n_req))
    who = rng.integers(0, len(sessions), n_req)
    lock = threading.Lock()
    tickets = {}          # ticket -> (session, arrival)
    lat, queue_ms, codes, iters, bounds = [], [], [], [], []
    finish64 = 0
    pool = cf.ThreadPoolExecutor(max_workers=workers)
    errors = []

    host = {"prepare": 0.0, "complete": 0.0}

    def start(s: Session):
        # session s is marked busy and has a pending arrival
        try:
            a = s.pending.popleft()
            t1 = time.perf_counter()
            s.prepare()
            host["prepare"] += time.perf_counter() - t1
            t = engine.submit(s.qp, solve_params(s, tol, precision, max_it=max_it))
            with lock:
                tickets[t] = (s, a)
        except Exception as e:  # surfaced by the main loop
            errors.append(e)

    def finish(s: Session, rc, st):
        try:
            t1 = time.perf_counter()
            keep = keep_history if isinstance(keep_history, bool) else s.sid in keep_history
            s.complete(keep, rc, st)
            host["complete"] += time.perf_counter() - t1
            with lock:
                again = bool(s.pending)
                if not again:
                    s.busy = False
            if again:
                start(s)
        except Exception as e:
            errors.append(e)

    t0 = time.perf_counter()
    k = 0
    done = 0
    while done < n_req:
        if errors:
            raise errors[0]
        now = time.perf_counter() - t0
        while k < n_req and arr[k] <= now:
            s = sessions[int(who[k])]
            with lock:
                s.pending.append(arr[k])
                go = not s.busy
                if go:
                    s.busy = True
            if go:
                pool.submit(start, s)
            k += 1
        wait_ms = 5 if k >= n_req else max(0, int((arr[k] - (time.perf_counter() - t0)) * 1e3))
        r = engine.wait_any(min(wait_ms, 5))
        if r is None:
            continue
        t, rc, st = r
        tdone = time.perf_counter() - t0
        while True:  # the submitting worker records the ticket right after submit returns
            with lock:
                entry = tickets.pop(t, None)
            if entry is not None:
                break
            time.sleep(1e-4)
        s, a = entry
        lat.append(1e3 * (tdone - a))
        queue_ms.append(st["queue_ms"])
        codes.append(rc)
        iters.append(st["iterations"])
        bounds.append(st["error_bound"])
        finish64 += st["fp64_finish"]
        done += 1
        pool.submit(finish, s, rc, st)
    wall = time.perf_counter() - t0
    pool.shutdown(wait=True)
    if errors:
        raise errors[0]
    lat = np.array(lat)
    return {"requests": n_req, "rate": rate, "wall_s": round(wall, 3),
            "solves_per_s": round(n_req / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "queue_p99_ms": round(float(np.percentile(queue_ms, 99)), 3),
            "converged": int(sum(1 for c in codes if c == 0)),
            "fp64_finish": int(finish64),
            "iterations_mean": round(float(np.mean(iters)), 1),
            "error_bound_m": {"max": float(np.max(bounds)), "median": float(np.median(bounds)),
                              "reported": int(sum(1 for b in bounds if b >= 0.0))},
            "host_prepare_ms": round(1e3 * host["prepare"] / n_req, 3),
            "host_complete_ms": round(1e3 * host["complete"] / n_req, 3),
            "engine": engine.info()}


def fp64_error_sample(engine: M.Engine, sessions, n=16, tol=1e-6, seed=8, precision=M.MIXED,
                      ref_tol=1e-10, max_it=1000):
    """The error of served solves against FP64 solves of the same QPs.
    For `n` sessions drawn at random, one step each: the engine's solve in
    `precision`, then, from the same warm start, the plain FP64 solve to
    `ref_tol` (the reference solution); the session then continues from the
    served solution.  Reports the relative and absolute errors and checks the
    served solve's error bound against the measured error (both solutions
    are within their bounds of the exact minimiser, so their distance is at
    most the sum of the two bounds)."""
    rng = np.random.default_rng(seed)
    picks = rng.choice(len(sessions), size=min(n, len(sessions)), replace=False)
    rel, ab, bound_ok, bounds = [], [], [], []
    for i in picks:
        s = sessions[int(i)]
        s.prepare()
        p = solve_params(s, tol, precision, max_it=max_it)
        rc, st = engine.wait(engine.submit(s.qp, p))
        x = s.qp.get_x()
        s.qp.reset_x()
        pr = solve_params(s, ref_tol, M.FP64, max_it=20 * max_it)
        pr.alpha_bar = st["alpha_bar"]
        rc64, st64 = s.qp.solve(pr)
        x64 = s.qp.get_x()
        s.qp.set_x(x)  # the session continues from the served solution
        d = float(np.linalg.norm(x - x64))
        rel.append(d / float(np.linalg.norm(x64)))
        ab.append(d)
        bounds.append(st["error_bound"])
        bound_ok.append(st["error_bound"] >= 0.0 and d <= st["error_bound"] + max(st64["error_bound"], 0.0))
        s.complete(rc=rc, stats=st)
    return {"sample": len(picks), "rel_err_max": float(np.max(rel)), "rel_err_median": float(np.median(rel)),
            "abs_err_max_m": float(np.max(ab)), "error_bound_max_m": float(np.max(bounds)),
            "bound_holds": bool(all(bound_ok)), "reference": f"plain FP64 solve to {ref_tol:g}"}


def run_one_at_a_time(sessions, n_req=200, tol=1e-6, seed=6, max_it=1000):
    """The plain FP64 form serving requests back to back (closed loop: the
    next request starts when the previous one ends); latency = prepare + solve."""
    rng = np.random.default_rng(seed)
    who = rng.integers(0, len(sessions), n_req)
    lat, codes = [], []
    t0 = time.perf_counter()
    for i in who:
        s = sessions[int(i)]
        a = time.perf_counter()
        s.prepare()
        rc, st = s.qp.solve(solve_params(s, tol, M.FP64, max_it=max_it))
        lat.append(1e3 * (time.perf_counter() - a))
        codes.append(rc)
        s.complete(rc=rc, stats=st)
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    return {"requests": n_req, "wall_s": round(wall, 3), "solves_per_s": round(n_req / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "converged": int(sum(1 for c in codes if c == 0))}


def main(argv=None):
    """Runs the serving workload and prints one JSON line."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--rate", type=float, default=2000.0, help="aggregate requests/s (Poisson)")
    ap.add_argument("--duration", type=float, default=5.0, help="seconds of arrivals")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--plain-requests", type=int, default=200)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--max-it", type=int, default=1000)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    ctx = M.Context(0)
    ss = make_sessions(ctx, count=a.sessions, seed=a.seed, own_contexts=True)
    dof = [3 * s.n_nodes for s in ss]
    print(f"setup: {len(ss)} sessions, DoF {min(dof)}..{max(dof)} (total {sum(dof)}), "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = M.Engine(ctx, max_active=256, tiles=4, hbm_budget=int(a.budget_gb * 1e9))
    prec = M.MIXED if a.precision == "mixed" else M.FP64
    run_open_loop(eng, ss, rate=a.rate, duration=0.3, tol=a.tol, precision=prec, seed=a.seed + 100,
                  workers=a.workers, max_it=a.max_it)  # warm-up: graph captures, FP32 copies
    res = run_open_loop(eng, ss, rate=a.rate, duration=a.duration, tol=a.tol, precision=prec,
                        seed=a.seed + 1, workers=a.workers, max_it=a.max_it)
    print(f"engine: {res}", file=sys.stderr, flush=True)
    plain = run_one_at_a_time(ss, n_req=a.plain_requests, tol=a.tol, seed=a.seed + 2, max_it=a.max_it)
    print(f"plain fp64 one at a time: {plain}", file=sys.stderr, flush=True)
    out = {"workload": f"{len(ss)} sessions, DoF {min(dof)}-{max(dof)}, Poisson {a.rate:g}/s",
           "precision": a.precision, "budget_gb": a.budget_gb, "engine": res, "plain_fp64": plain}
    print(json.dumps(out), flush=True)
    eng.close()
    ctx.close()
    return out


if __name__ == "__main__":
    main()
