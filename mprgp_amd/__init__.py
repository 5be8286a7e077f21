"""MI355X-native MPRGP contact-QP solver (host side over the C ABI).

Object API (device-resident, mirrors include/mprgp_amd.h)::

    ctx = Context(0)
    A = ctx.matrix_csr(A_scipy)            # or ctx.matrix_bsr3(n, rowptr, colidx, blocks)
    qp = A.qp()
    qp.set_rows(node, normal, c); qp.set_rhs(b); qp.set_x(x0)
    code, stats = qp.solve(tol=1e-8, max_it=500)
    x = qp.get_x()

Reference-shaped facade (``MPRGPDecoupledCon<double>::solve``,
src/MPRGPSolver/MPRGPSolver.h:538-590)::

    code = MPRGPDecoupledCon.solve(A, B, J, c, x, tol, max_it, "MPRGP", ev)

Every call runs the HIP kernels of ``libmprgp_amd.so``; without the built
library (or without a GPU) the calls raise instead of falling back.
"""
from __future__ import annotations

import ctypes as C
import os
from collections.abc import Sequence

import numpy as np

from . import _lib
from ._lib import (CLASSIC, COROTATIONAL, FP64, LINEAR, MIXED, MONOTONIC, RHS_ONLY, MprgpError,
                   Params, Stats, StepParams, StepStats, check, last_error, load)

__all__ = ["Context", "Matrix", "QP", "Fem", "Engine", "MPRGPDecoupledCon", "solve_batch", "params",
           "read_qp", "write_qp", "qp_file_info",
           "MprgpError", "MONOTONIC", "CLASSIC", "FP64", "MIXED", "LINEAR", "COROTATIONAL",
           "RHS_ONLY", "device_count"]


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _planes(planes):
    """(K, 5) float64 {nx, ny, nz, p, delta} from an array or (normal, p, delta) tuples."""
    if isinstance(planes, np.ndarray):
        return _f64(planes).reshape(-1, 5)
    planes = list(planes)
    if all(len(pl) == 5 and all(np.isscalar(v) for v in pl) for pl in planes):
        return _f64(np.array(planes, np.float64).reshape(-1, 5))
    rows = [[*np.asarray(n, np.float64).ravel(), float(p), float(d)] for n, p, d in planes]
    return _f64(np.array(rows, np.float64).reshape(-1, 5))


def device_count() -> int:
    n = C.c_int(0)
    rc = load().mprgp_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def params(**kw) -> Params:
    p = Params()
    load().mprgp_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Context:
    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = C.c_void_p()
        check(load().mprgp_ctx_create(int(device), C.c_size_t(hbm_budget), C.byref(self.h)))
        self._objs = []

    def close(self):
        if self.h:
            # engines first: they may still hold QPs
            for o in [o for o in self._objs if isinstance(o, Engine)] + \
                     [o for o in reversed(self._objs) if not isinstance(o, Engine)]:
                o.close()
            self._objs = []
            load().mprgp_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        check(load().mprgp_ctx_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def copy_probe(self, nbytes=1 << 30, iters=20) -> float:
        """Device copy rate in GB/s (read + write) of a streaming kernel."""
        g = C.c_double()
        check(load().mprgp_copy_probe(self.h, C.c_size_t(int(nbytes)), int(iters), C.byref(g)))
        return g.value

    def matrix_bsr3(self, n_nodes, rowptr, colidx, blocks) -> "Matrix":
        rowptr, colidx = _i32(rowptr), _i32(colidx)
        blocks = _f64(blocks)
        h = C.c_void_p()
        check(load().mprgp_matrix_create_bsr3(self.h, int(n_nodes), _p(rowptr), _p(colidx),
                                              _p(blocks), C.byref(h)))
        m = Matrix(self, h, int(n_nodes))
        self._objs.append(m)
        return m

    def matrix_csr(self, A) -> "Matrix":
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        ptr, idx, val = _i32(A.indptr), _i32(A.indices), _f64(A.data)
        h = C.c_void_p()
        check(load().mprgp_matrix_create_csr(self.h, int(A.shape[0]), _p(ptr), _p(idx), _p(val),
                                             C.byref(h)))
        m = Matrix(self, h, A.shape[0] // 3)
        self._objs.append(m)
        return m


class Fem:
    """Device FEM assembly of one body (buildLinearSystem) and, once a state
    is set, its implicit-Euler time step (MprgpFemSolver::advance) with the
    positions and velocities kept on the device."""

    def __init__(self, ctx: Context, rest_pos, tets, young, poisson, density):
        X = _f64(rest_pos)
        T = _i32(tets)
        nt = T.size // 4
        E = _f64(np.broadcast_to(np.asarray(young, np.float64), (nt,)))
        nu = _f64(np.broadcast_to(np.asarray(poisson, np.float64), (nt,)))
        self.h = C.c_void_p()
        check(load().mprgp_fem_create(ctx.h, X.size // 3, nt, _p(T), _p(X), _p(E), _p(nu),
                                      float(density), C.byref(self.h)))
        self.ctx = ctx
        self.n_nodes = X.size // 3
        self.matrix = Matrix(ctx, C.c_void_p(load().mprgp_fem_matrix(self.h)), self.n_nodes)
        self.matrix.owned_by_fem = True
        ctx._objs.append(self)

    def close(self):
        if self.h:
            load().mprgp_fem_destroy(self.h)
            self.h = C.c_void_p()
            self.matrix.h = C.c_void_p()

    def assemble(self, qp: "QP", x_n, v_n, gravity=(0.0, -9.8, 0.0), h=0.01, material=LINEAR):
        x = _f64(x_n, 3 * self.n_nodes)
        v = _f64(v_n, 3 * self.n_nodes)
        g = _f64(gravity, 3)
        check(load().mprgp_fem_assemble(self.h, int(material), _p(x), _p(v), _p(g), float(h), qp.h))

    def assemble_at(self, qp: "QP", x_n, v_n, x1, gravity=(0.0, -9.8, 0.0), h=0.01, material=LINEAR):
        """A and b at the Newton iterate x1 (x1 = None: x_n)."""
        x = _f64(x_n, 3 * self.n_nodes)
        v = _f64(v_n, 3 * self.n_nodes)
        y = None if x1 is None else _f64(x1, 3 * self.n_nodes)
        g = _f64(gravity, 3)
        check(load().mprgp_fem_assemble_at(self.h, int(material), _p(x), _p(v), None if y is None else _p(y),
                                           _p(g), float(h), qp.h))

    def set_damping(self, alpha_k=0.0, alpha_m=0.0):
        """Rayleigh damping: mass term alpha_k, stiffness term alpha_m (the scene's keys)."""
        check(load().mprgp_fem_set_damping(self.h, float(alpha_k), float(alpha_m)))

    def set_state(self, x, v):
        check(load().mprgp_fem_set_state(self.h, _p(_f64(x, 3 * self.n_nodes)),
                                         _p(_f64(v, 3 * self.n_nodes))))

    def get_state(self):
        x = np.empty(3 * self.n_nodes)
        v = np.empty(3 * self.n_nodes)
        check(load().mprgp_fem_get_state(self.h, _p(x), _p(v)))
        return x, v

    def step_params(self, h=0.01, gravity=(0.0, -9.8, 0.0), material=LINEAR, planes=(), newton_max_it=10,
                    newton_eps=1e-4):
        """StepParams (the planes array is kept alive on the returned object)."""
        sp = StepParams()
        sp.h = float(h)
        for k in range(3):
            sp.gravity[k] = float(gravity[k])
        sp.material = int(material)
        P = _planes(planes)
        sp._planes = P
        sp.n_planes = int(P.shape[0])
        sp.planes = P.ctypes.data if P.size else None
        sp.newton_max_it = int(newton_max_it)
        sp.newton_eps = float(newton_eps)
        return sp

    def advance(self, qp: "QP", sp: StepParams, p: Params | None = None, **kw):
        """One time step; returns (code, step stats dict)."""
        p = params(**kw) if p is None else p
        st = StepStats()
        rc = load().mprgp_fem_advance(self.h, qp.h, C.byref(sp), C.byref(p), C.byref(st))
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED))
        return rc, st.as_dict()

    def begin_step(self, qp: "QP", sp: StepParams):
        rows = C.c_int(0)
        check(load().mprgp_fem_begin_step(self.h, qp.h, C.byref(sp), C.byref(rows)))
        return rows.value

    def end_step(self, qp: "QP", sp: StepParams):
        delta = C.c_double(0.0)
        check(load().mprgp_fem_end_step(self.h, qp.h, C.byref(sp), C.byref(delta)))
        return delta.value


class Matrix:
    def __init__(self, ctx, h, n_nodes):
        self.ctx, self.h, self.n_nodes = ctx, h, n_nodes

    owned_by_fem = False

    def close(self):
        if self.h and not self.owned_by_fem:
            load().mprgp_matrix_destroy(self.h)
        self.h = C.c_void_p()

    def info(self):
        n, nb, by = C.c_int(), C.c_longlong(), C.c_longlong()
        check(load().mprgp_matrix_info(self.h, C.byref(n), C.byref(nb), C.byref(by)))
        return {"n_nodes": n.value, "n_blocks": nb.value, "bytes": by.value}

    def update_bsr3(self, blocks):
        b = _f64(blocks)
        check(load().mprgp_matrix_update_bsr3(self.h, _p(b)))

    def spmv(self, x):
        x = _f64(x, 3 * self.n_nodes)
        y = np.empty(3 * self.n_nodes)
        check(load().mprgp_matrix_spmv(self.h, _p(x), _p(y)))
        return y

    def func_value(self, x, b):
        """1/2 x'Ax - x'b (FixedSparseMatrix::funcValue)."""
        f = C.c_double(0.0)
        check(load().mprgp_matrix_func_value(self.h, _p(_f64(x, 3 * self.n_nodes)),
                                             _p(_f64(b, 3 * self.n_nodes)), C.byref(f)))
        return f.value

    def diagonal(self):
        d = np.empty(3 * self.n_nodes)
        check(load().mprgp_matrix_diagonal(self.h, _p(d)))
        return d

    def get_bsr3(self):
        """(rowptr, colidx, blocks[nb, 3, 3]) in the creation's block order."""
        i = self.info()
        rp = np.zeros(self.n_nodes + 1, np.int32)
        ci = np.zeros(i["n_blocks"], np.int32)
        bl = np.zeros((i["n_blocks"], 3, 3))
        check(load().mprgp_matrix_get_bsr3(self.h, _p(rp), _p(ci), _p(bl)))
        return rp, ci, bl

    def to_csr(self):
        """The matrix as a scipy CSR of scalars (3 rows per node)."""
        import scipy.sparse as sp
        rp, ci, bl = self.get_bsr3()
        return sp.bsr_matrix((bl, ci, rp), shape=(3 * self.n_nodes, 3 * self.n_nodes)).tocsr()

    def set_lambda_min(self, lower: float):
        """A lower bound on the smallest eigenvalue (0: none); solves then
        report error_bound = residual / lower."""
        check(load().mprgp_matrix_set_lambda_min(self.h, float(lower)))

    def qp(self) -> "QP":
        h = C.c_void_p()
        check(load().mprgp_qp_create(self.h, C.byref(h)))
        q = QP(self, h)
        self.ctx._objs.append(q)
        return q


class QP:
    """Device-resident QP: matrix + half-space rows + b + x (+ power vector)."""

    def __init__(self, A: Matrix, h):
        self.A, self.h = A, h
        self.n = 3 * A.n_nodes

    def close(self):
        if self.h:
            load().mprgp_qp_destroy(self.h)
            self.h = C.c_void_p()

    def set_rows(self, node, normal, c):
        node = _i32(node)
        m = node.size
        normal = _f64(normal, 3 * m)
        c = _f64(c, m)
        check(load().mprgp_qp_set_rows(self.h, int(m), _p(node), _p(normal), _p(c)))

    def set_rows_sparse(self, J, c):
        """J: scipy sparse m x n (rows inside one node each)."""
        import scipy.sparse as sp
        J = sp.csc_matrix(J)
        J.sort_indices()
        c = _f64(c, J.shape[0])
        outer, inner, val = _i32(J.indptr), _i32(J.indices), _f64(J.data)
        check(load().mprgp_qp_set_rows_sparse(self.h, int(J.shape[0]), 1, _p(outer), _p(inner),
                                              _p(val), _p(c)))

    def collide_planes(self, planes):
        """Rows of the decoupled plane collision at the QP's current x, made on
        the device (LinearConCollider::handle per plane, first plane wins).
        planes: (K, 5) {nx, ny, nz, p, delta} or a list of (normal, p, delta)
        with unit normals.  Returns the row count."""
        P = _planes(planes)
        m = C.c_int(0)
        check(load().mprgp_qp_collide_planes(self.h, int(P.shape[0]), _p(P), C.byref(m)))
        return m.value

    def get_rows(self):
        """(node, normal (m, 3), c) of the QP's rows, by ascending node."""
        m = C.c_int(0)
        L = load()
        check(L.mprgp_qp_get_rows(self.h, 0, C.byref(m), None, None, None))
        node = np.zeros(max(m.value, 1), np.int32)
        normal = np.zeros((max(m.value, 1), 3))
        c = np.zeros(max(m.value, 1))
        check(L.mprgp_qp_get_rows(self.h, m.value, C.byref(m), _p(node), _p(normal), _p(c)))
        k = m.value
        return node[:k], normal[:k], c[:k]

    # DecoupledConProjector / DiagonalDecouplePreconSolver on this QP's rows
    def project(self, x):
        y = np.empty(self.n)
        check(load().mprgp_qp_project(self.h, _p(_f64(x, self.n)), _p(y)))
        return y

    def is_feasible(self, x):
        ok = C.c_int(0)
        check(load().mprgp_qp_is_feasible(self.h, _p(_f64(x, self.n)), C.byref(ok)))
        return bool(ok.value)

    def decide_face(self, x):
        """Sets the QP's face from x; returns the active count."""
        k = C.c_int(0)
        check(load().mprgp_qp_decide_face(self.h, _p(_f64(x, self.n)), C.byref(k)))
        return k.value

    def phi(self, g):
        out = np.empty(self.n)
        check(load().mprgp_qp_phi(self.h, _p(_f64(g, self.n)), _p(out)))
        return out

    def beta(self, g, phi):
        out = np.empty(self.n)
        check(load().mprgp_qp_beta(self.h, _p(_f64(g, self.n)), _p(_f64(phi, self.n)), _p(out)))
        return out

    def phitphi(self, x, alpha_bar, phi):
        out = C.c_double(0.0)
        check(load().mprgp_qp_phitphi(self.h, _p(_f64(x, self.n)), float(alpha_bar), _p(_f64(phi, self.n)),
                                      C.byref(out)))
        return out.value

    def step_limit(self, x, d, alpha_cg=np.finfo(np.float64).max):
        out = C.c_double(0.0)
        check(load().mprgp_qp_step_limit(self.h, _p(_f64(x, self.n)), _p(_f64(d, self.n)), float(alpha_cg),
                                         C.byref(out)))
        return out.value

    def precond(self, g, phi):
        out = np.empty(self.n)
        check(load().mprgp_qp_precond(self.h, _p(_f64(g, self.n)), _p(_f64(phi, self.n)), _p(out)))
        return out

    def set_rhs(self, b):
        b = _f64(b, self.n)
        check(load().mprgp_qp_set_rhs(self.h, _p(b)))

    def get_rhs(self):
        b = np.empty(self.n)
        check(load().mprgp_qp_get_rhs(self.h, _p(b)))
        return b

    def set_x(self, x):
        x = _f64(x, self.n)
        check(load().mprgp_qp_set_x(self.h, _p(x)))

    def reset_x(self):
        check(load().mprgp_qp_reset_x(self.h))

    def get_x(self):
        x = np.empty(self.n)
        check(load().mprgp_qp_get_x(self.h, _p(x)))
        return x

    def set_ev(self, ev):
        if ev is None:
            check(load().mprgp_qp_set_ev(self.h, None))
        else:
            ev = _f64(ev, self.n)
            check(load().mprgp_qp_set_ev(self.h, _p(ev)))

    def get_ev(self):
        ev = np.empty(self.n)
        v = C.c_int()
        check(load().mprgp_qp_get_ev(self.h, _p(ev), C.byref(v)))
        return ev if v.value else None

    def op_counts(self):
        c = np.zeros(32, dtype=np.int32)
        check(load().mprgp_qp_op_counts(self.h, _p(c), 32))
        return c

    def get_face(self):
        f = np.empty(self.n // 3, dtype=np.uint8)
        check(load().mprgp_qp_get_face(self.h, _p(f)))
        return f

    def solve(self, p: Params | None = None, **kw):
        p = params(**kw) if p is None else p
        st = Stats()
        rc = load().mprgp_qp_solve(self.h, C.byref(p), C.byref(st))
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
        return rc, st.as_dict()


def general_solve(qp, J, c, p: Params | None = None, **kw):
    """MPRGPGeneralCon::solve on a QP's matrix, b and x with coupled rows
    J x >= c (J: scipy sparse, m x 3*nodes, user DoF order).  Returns (rc, stats)."""
    import scipy.sparse as sp
    p = params(**kw) if p is None else p
    Jc = sp.csr_matrix(J)
    Jc.sort_indices()
    jp, jc_, jv = _i32(Jc.indptr), _i32(Jc.indices), _f64(Jc.data)
    c = _f64(c, Jc.shape[0])
    st = Stats()
    rc = load().mprgp_general_solve(qp.h, int(Jc.shape[0]), _p(jp), _p(jc_), _p(jv), _p(c), C.byref(p),
                                    C.byref(st))
    check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
    return rc, st.as_dict()


def ica_solve(qp, J, p, tol=1e-4, max_it=1000, gs_max_it=20, gs_tol=1e-3):
    """ICASolver(max_it, tol, gs_max_it, gs_tol).reset(A, b).solve(J, p, x)
    (ICASolver.h:287-415) on a QP's matrix, b and x (x overwritten), rows
    J x >= p (scipy sparse, user DoF order).  Returns (rc, stats, lambda,
    projected Gauss-Seidel sweeps)."""
    import scipy.sparse as sp
    Jc = sp.csr_matrix(J)
    Jc.sort_indices()
    m = Jc.shape[0]
    jp, jc_, jv = _i32(Jc.indptr), _i32(Jc.indices), _f64(Jc.data)
    pv = _f64(p, m) if m else np.zeros(1)
    lam = np.zeros(max(m, 1))
    sweeps = C.c_int(0)
    st = Stats()
    rc = load().mprgp_ica_solve(qp.h, int(m), _p(jp), _p(jc_), _p(jv), _p(pv), float(tol), int(max_it),
                                int(gs_max_it), float(gs_tol), _p(lam), C.byref(sweeps), C.byref(st))
    check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED))
    return rc, st.as_dict(), lam[:m], sweeps.value


class BatchStats(Sequence):
    """The per-QP stats of a batch solve as a sequence of dicts, each made
    from the returned C array on first access."""

    __slots__ = ("_raw", "_dicts")

    def __init__(self, raw):
        self._raw = raw
        self._dicts = [None] * len(raw)

    def __len__(self):
        return len(self._dicts)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(len(self)))]
        n = len(self._dicts)
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError("batch stats index out of range")
        d = self._dicts[i]
        if d is None:
            d = self._dicts[i] = self._raw[i].as_dict()
        return d

    def __repr__(self):
        return repr(list(self))


def solve_batch(ctx: Context, qps, p=None, per_qp=None, **kw):
    """Advances all QPs together; returns (rc, stats): a sequence of one
    stats dict per QP."""
    n = len(qps)
    arr = (C.c_void_p * n)(*[q.h.value for q in qps])
    if per_qp is not None:
        ps = (Params * n)(*per_qp)
        stride = 1
    else:
        ps = (Params * 1)(p if p is not None else params(**kw))
        stride = 0
    st = (Stats * n)()
    rc = load().mprgp_solve_batch(ctx.h, n, C.cast(arr, C.POINTER(C.c_void_p)), ps, stride, st)
    check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
    return rc, BatchStats(st)


class Engine:
    """Continuous-batching solve engine (``mprgp_engine_*``): ``submit`` queues a
    device-resident QP and returns a ticket at once; a native worker thread
    admits it into the running batch between graph replays (slots and the HBM
    budget permitting); ``wait`` returns its code and stats (``elapsed_ms`` is
    the submit-to-completion latency, ``queue_ms`` the wait for admission)."""

    def __init__(self, ctx: Context, max_active=256, tiles=4, graph_steps=8, hbm_budget=0):
        cfg = _lib.EngineConfig()
        load().mprgp_engine_config_default(C.byref(cfg))
        cfg.max_active, cfg.tiles, cfg.graph_steps = int(max_active), int(tiles), int(graph_steps)
        cfg.hbm_budget = int(hbm_budget)
        self.h = C.c_void_p()
        check(load().mprgp_engine_create(ctx.h, C.byref(cfg), C.byref(self.h)))
        self.ctx = ctx
        ctx._objs.append(self)

    def close(self):
        if self.h:
            load().mprgp_engine_destroy(self.h)
            self.h = C.c_void_p()

    def submit(self, qp: "QP", p: Params | None = None, **kw) -> int:
        p = params(**kw) if p is None else p
        t = C.c_longlong()
        check(load().mprgp_engine_submit(self.h, qp.h, C.byref(p), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """Returns (rc, stats) or None when the ticket is still running after timeout_ms."""
        st = Stats()
        rc = load().mprgp_engine_wait(self.h, C.c_longlong(ticket), int(timeout_ms), C.byref(st))
        if rc == _lib.EPENDING:
            return None
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
        d = st.as_dict()
        return rc, d

    def wait_any(self, timeout_ms: int = -1):
        """Returns (ticket, rc, stats) of the first completed ticket, or None on time-out."""
        st = Stats()
        t = C.c_longlong()
        rc = load().mprgp_engine_wait_any(self.h, int(timeout_ms), C.byref(t), C.byref(st))
        if rc == _lib.EPENDING:
            return None
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
        d = st.as_dict()
        return t.value, rc, d

    def info(self):
        i = _lib.EngineInfo()
        check(load().mprgp_engine_get_info(self.h, C.byref(i)))
        return i.as_dict()


class MPRGPDecoupledCon:
    """Mirror of ``MPRGPDecoupledCon<double>`` (MPRGPSolver.h:538-590)."""

    @staticmethod
    def solve(A, B, J, c, x, tol=1e-3, max_it=1000, solver_name="MPRGP", ev=None,
              device=0, stats_out=None):
        """J/c overload (MPRGPSolver.h:561-574): projects x first, solves with the
        tangent-Jacobi preconditioner, writes the solution into ``x`` (in place) and
        returns the reference's code (0 converged, -1 not).  ``ev`` (a numpy array of
        length n, or None) is the power-method vector carried across solves."""
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        Jc = sp.csc_matrix(J, shape=(J.shape[0], A.shape[0]))
        Jc.sort_indices()
        n = A.shape[0]
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
            raise TypeError("x must be a contiguous float64 numpy array (it is overwritten)")
        B = _f64(B, n)
        c = _f64(c, Jc.shape[0])
        ap, ai, av = _i32(A.indptr), _i32(A.indices), _f64(A.data)
        jp, ji, jv = _i32(Jc.indptr), _i32(Jc.indices), _f64(Jc.data)
        evb = np.zeros(n) if ev is None else ev
        valid = C.c_int(0 if ev is None or not np.any(ev) else 1)
        st = Stats()
        rc = load().mprgp_decoupled_solve(int(device), n, _p(ap), _p(ai), _p(av), _p(B),
                                          int(Jc.shape[0]), _p(jp), _p(ji), _p(jv), _p(c), _p(x),
                                          float(tol), int(max_it), _p(evb), C.byref(valid),
                                          C.byref(st))
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
        if stats_out is not None:
            stats_out.update(st.as_dict())
        return 0 if rc == _lib.OK else -1

    @staticmethod
    def solve_file(file_name, tol=1e-3, max_it=1000, solver_name="MPRGP", ev=None, device=0,
                   stats_out=None):
        """File overload (MPRGPSolver.h:576-589): reads A, B, J, c and x0 from a
        QP file in the reference's binary layout (loadQP, MPRGPUtility.h:446-487)
        and solves from x0 as ``solve``.  The reference fills its x argument;
        here the solution is returned: ``(code, x)``."""
        n, m, _, _ = qp_file_info(file_name)
        x = np.zeros(n)
        evb = np.zeros(n) if ev is None else _f64(ev, n)
        valid = C.c_int(0 if ev is None or not np.any(ev) else 1)
        st = Stats()
        n_out = C.c_int(0)
        rc = load().mprgp_decoupled_solve_file(int(device), os.fsencode(file_name), _p(x), n,
                                               C.byref(n_out), float(tol), int(max_it), _p(evb),
                                               C.byref(valid), C.byref(st))
        check(rc, allow=(_lib.OK, _lib.NOT_CONVERGED, _lib.ENAN))
        if ev is not None and isinstance(ev, np.ndarray) and ev.dtype == np.float64 and ev.size == n:
            ev[:] = evb
        if stats_out is not None:
            stats_out.update(st.as_dict())
        return (0 if rc == _lib.OK else -1), x


def qp_file_info(file_name):
    """(n, m, nnz of A, nnz of J) of a QP file (duplicates summed)."""
    n, m = C.c_int(), C.c_int()
    an, jn = C.c_longlong(), C.c_longlong()
    check(load().mprgp_qp_file_info(os.fsencode(file_name), C.byref(n), C.byref(m), C.byref(an),
                                    C.byref(jn)))
    return n.value, m.value, an.value, jn.value


def read_qp(file_name):
    """(A, B, J, c, x0) of a QP file in the reference's binary layout, A and J
    as scipy CSR matrices (loadQP, MPRGPUtility.h:446-487)."""
    import scipy.sparse as sp
    n, m, an, jn = qp_file_info(file_name)
    ap, ac, av = np.zeros(n + 1, np.int32), np.zeros(an, np.int32), np.zeros(an)
    jp, jc, jv = np.zeros(m + 1, np.int32), np.zeros(jn, np.int32), np.zeros(jn)
    B, c, x0 = np.zeros(n), np.zeros(m), np.zeros(n)
    check(load().mprgp_qp_file_read(os.fsencode(file_name), _p(ap), _p(ac), _p(av), _p(B), _p(jp),
                                    _p(jc), _p(jv), _p(c), _p(x0)))
    return (sp.csr_matrix((av, ac, ap), shape=(n, n)), B, sp.csr_matrix((jv, jc, jp), shape=(m, n)),
            c, x0)


def write_qp(file_name, A, B, J, c, x0):
    """writeQP (MPRGPUtility.h:406-443): A (n x n), J (m x n) scipy sparse."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    n = A.shape[0]
    J = sp.csr_matrix(J, shape=(J.shape[0], n))
    m = J.shape[0]
    ap, ac, av = _i32(A.indptr), _i32(A.indices), _f64(A.data)
    jp, jc, jv = _i32(J.indptr), _i32(J.indices), _f64(J.data)
    B, c, x0 = _f64(B, n), _f64(c, m), _f64(x0, n)
    check(load().mprgp_write_qp(os.fsencode(file_name), n, _p(ap), _p(ac), _p(av), _p(B), m, _p(jp),
                                _p(jc), _p(jv), _p(c), _p(x0)))
