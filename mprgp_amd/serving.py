"""Serving loop: simulation sessions issuing one contact QP per timestep to
the continuous-batching engine.

A session is one deformable body (a beam or a blob, tet mesh) on seeded
obstacle planes.  Its matrix, right-hand side, rows and warm start live in one
device-resident QP.  A timestep (one request) is

* every ``reassemble`` steps: a random rotation of the body about its
  centroid, full corotational assembly of A and b at x_n, and the collider's
  rows recomputed from x_n (host, as the reference's collision handling);
* otherwise: right-hand side only (``RHS_ONLY``: the lagged matrix is kept);
* the solve from the warm start x_n (projected first, as
  DecoupledMprgpFemSolver::forward does, MprgpFemSolver.cpp:528-533);
* x_{n+1} = x, v_{n+1} = (x - x_n) / h.

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

    @property
    def n_nodes(self):
        return self.scene.n_nodes

    def prepare(self):
        """Host + device work before the solve of the current step."""
        sc = self.scene
        full = self.step % self.reassemble == 0
        if full and self.step > 0:
            ang = math.radians(self.rng.uniform(1.0, 10.0))
            axis = self.rng.normal(size=3)
            axis /= np.linalg.norm(axis)
            K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
            R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K)
            c = sc.x_n.mean(axis=0)
            sc.x_n = (sc.x_n - c) @ R.T + c
            sc.v_n = sc.v_n @ R.T
        material = M.COROTATIONAL if full else (M.COROTATIONAL | M.RHS_ONLY)
        self.fem.assemble(self.qp, sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, material)
        if full:
            node, normal, c = sc.rows()
            self.qp.set_rows(node, normal, c)
        self.qp.set_x(sc.x_n.ravel())

    def complete(self, keep_history=False):
        sc = self.scene
        x = self.qp.get_x().reshape(-1, 3)
        if keep_history:
            self.history.append(x.copy())
        sc.v_n = (x - sc.x_n) / sc.h
        sc.x_n = x
        self.step += 1


def session_scene(rng, sid, dof_range=(3e3, 1.5e5)):
    """Seeded beam or blob with DoF log-uniform in dof_range, E log-uniform
    [1e4, 1e6], nu uniform [0.3, 0.48], 1-3 planes placing 2-10% of the nodes
    on them (the ground is always one)."""
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
    k = int(rng.integers(1, 4))
    frac = rng.uniform(0.02, 0.10)
    sc.planes = [((0.0, 1.0, 0.0), 0.0, 0.01)] + problems.random_planes(rng, sc.x_n, k - 1, frac) \
        if k > 1 else [((0.0, 1.0, 0.0), 0.0, 0.01)]
    return sc


def make_sessions(ctx, count=64, seed=5, dof_range=(3e3, 1.5e5), reassemble=10):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        sc = session_scene(rng, i, dof_range)
        fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
        qp = fem.matrix.qp()
        out.append(Session(i, sc, fem, qp, reassemble=reassemble,
                           rng=np.random.default_rng([seed, i])))
    # alpha_bar per session from its first matrix (power method, batched)
    for s in out:
        s.prepare()
    rc, st = M.solve_batch(ctx, [s.qp for s in out], M.params(tol=1e-6, max_it=0, power_tol=1e-6))
    for s, d in zip(out, st):
        s.alpha_bar = d["alpha_bar"]
    return out


def solve_params(s: Session, tol, precision, tiles=4):
    return M.params(tol=tol, max_it=4000, alpha_bar=s.alpha_bar, project_first=1,
                    precision=precision, tiles=tiles)


def run_open_loop(engine: M.Engine, sessions, rate=2000.0, duration=3.0, tol=1e-6,
                  precision=M.MIXED, seed=6, workers=4, keep_history=False):
    """Poisson arrivals at `rate` requests/s for `duration` s, each to a
    uniformly chosen session; a session's requests are served in order.
    Latency = completion of the solve - arrival of the request."""
    rng = np.random.default_rng(seed)
    n_req = max(1, int(rng.poisson(rate * duration)))
    arr = np.sort(rng.uniform(0.0, duration, n_req))
    who = rng.integers(0, len(sessions), n_req)
    lock = threading.Lock()
    tickets = {}          # ticket -> (session, arrival)
    lat, queue_ms, codes, iters = [], [], [], []
    pool = cf.ThreadPoolExecutor(max_workers=workers)
    errors = []

    def start(s: Session):
        # session s is marked busy and has a pending arrival
        try:
            a = s.pending.popleft()
            s.prepare()
            t = engine.submit(s.qp, solve_params(s, tol, precision))
            with lock:
                tickets[t] = (s, a)
        except Exception as e:  # surfaced by the main loop
            errors.append(e)

    def finish(s: Session):
        try:
            s.complete(keep_history)
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
        done += 1
        pool.submit(finish, s)
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
            "iterations_mean": round(float(np.mean(iters)), 1),
            "engine": engine.info()}


def run_one_at_a_time(sessions, n_req=200, tol=1e-6, seed=6):
    """The plain FP64 form serving requests back to back (closed loop: the
    next request starts when the previous one ends); latency = prepare + solve."""
    rng = np.random.default_rng(seed)
    who = rng.integers(0, len(sessions), n_req)
    lat = []
    t0 = time.perf_counter()
    for i in who:
        s = sessions[int(i)]
        a = time.perf_counter()
        s.prepare()
        s.qp.solve(solve_params(s, tol, M.FP64))
        lat.append(1e3 * (time.perf_counter() - a))
        s.complete()
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    return {"requests": n_req, "wall_s": round(wall, 3), "solves_per_s": round(n_req / wall, 1),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3)}
