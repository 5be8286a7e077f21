"""CPU oracle for the MPRGP contact-QP path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/liboracle.so`` (built from ``mprgp_oracle.c`` and
``fem_oracle.c`` by ``oracle/Makefile``).  Only the tests, the smoke check and
the CPU-baseline leg of the benchmark import this package; the product library
never does.  The C sources state, function by function, which reference
file:line they restate and how each function is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

MONOTONIC = 0
CLASSIC = 1
LINEAR = 1
COROTATIONAL = 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    cmd = ["make", "-s", "-C", _HERE]
    if force:
        cmd.insert(2, "-B")
    subprocess.run(cmd, check=True)
    return _LIB_PATH


class _CSR(C.Structure):
    _fields_ = [("rows", C.c_int), ("cols", C.c_int), ("ptr", C.c_void_p),
                ("idx", C.c_void_p), ("val", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("code", C.c_int), ("iterations", C.c_int), ("num_cg", C.c_int),
                ("num_exp", C.c_int), ("num_prop", C.c_int), ("active", C.c_int),
                ("power_iters", C.c_int), ("residual", C.c_double),
                ("alpha_bar", C.c_double), ("lambda_max", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        D = C.c_double
        I = C.c_int
        L.orc_spmv.argtypes = [C.POINTER(_CSR), P, P]
        L.orc_func_value.argtypes = [C.POINTER(_CSR), P, P]
        L.orc_func_value.restype = D
        L.orc_projector_new.argtypes = [C.POINTER(_CSR), P]
        L.orc_projector_new.restype = P
        L.orc_projector_free.argtypes = [P]
        L.orc_projector_is_decoupled.argtypes = [P]
        L.orc_project.argtypes = [P, P, P]
        L.orc_decide_face.argtypes = [P, P]
        L.orc_is_feasible.argtypes = [P, P]
        L.orc_phi.argtypes = [P, P, P]
        L.orc_beta.argtypes = [P, P, P, P]
        L.orc_phitphi.argtypes = [P, P, D, P]
        L.orc_phitphi.restype = D
        L.orc_step_limit.argtypes = [P, P, P, D]
        L.orc_step_limit.restype = D
        L.orc_face.argtypes = [P]
        L.orc_face.restype = C.POINTER(C.c_char)
        L.orc_num_active.argtypes = [P]
        L.orc_set_face.argtypes = [P, P]
        L.orc_precond_new.argtypes = [C.POINTER(_CSR), P, I]
        L.orc_precond_new.restype = P
        L.orc_precond_free.argtypes = [P]
        L.orc_precond_solve.argtypes = [P, P, P, P, P]
        L.orc_spec_rad.argtypes = [C.POINTER(_CSR), P, P, D, C.c_ulonglong, P]
        L.orc_spec_rad.restype = D
        L.orc_decoupled_solve.argtypes = [C.POINTER(_CSR), P, C.POINTER(_CSR), P, P, D, I, I, I,
                                          D, P, P, I, C.POINTER(Stats)]
        L.orc_decoupled_solve_ex.argtypes = [C.POINTER(_CSR), P, C.POINTER(_CSR), P, P, D, I, I, I,
                                             D, P, P, I, C.POINTER(_Options), C.POINTER(Stats)]
        L.orc_qp_nonneg.argtypes = [I, P, P, P]
        L.orc_gprojector_new.argtypes = [C.POINTER(_CSR), P]
        L.orc_gprojector_new.restype = P
        L.orc_gprojector_free.argtypes = [P]
        L.orc_gproject.argtypes = [P, P, P]
        L.orc_gdecide_face.argtypes = [P, P]
        L.orc_gis_feasible.argtypes = [P, P]
        L.orc_gphi.argtypes = [P, P, P]
        L.orc_gbeta.argtypes = [P, P, P, P]
        L.orc_gphitphi.argtypes = [P, P, D, P]
        L.orc_gphitphi.restype = D
        L.orc_gface.argtypes = [P]
        L.orc_gface.restype = C.POINTER(C.c_char)
        L.orc_gnum_active.argtypes = [P]
        L.orc_general_solve.argtypes = [C.POINTER(_CSR), P, C.POINTER(_CSR), P, P, D, I, I, D, P, P, I,
                                        C.POINTER(Stats)]
        L.orc_inv3.argtypes = [P, P]
        L.orc_pgs.argtypes = [C.POINTER(_CSR), P, P, I, D, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_pgs.restype = I
        L.orc_ica_b.argtypes = [C.POINTER(_CSR), I, P, C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int)),
                                C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_double))]
        L.orc_ica_free.argtypes = [P]
        L.orc_ica_solve.argtypes = [C.POINTER(_CSR), P, C.POINTER(_CSR), P, P, D, I, I, D, P,
                                    C.POINTER(IcaStats)]
        L.orc_fem_setup.argtypes = [I, I, P, P, P, P, P]
        L.orc_polar_rotation.argtypes = [P, P]
        L.orc_fem_element.argtypes = [I, P, P, D, D, D, P, P]
        L.orc_fem_element.restype = D
        L.orc_fem_build_system.argtypes = [I, I, P, P, P, P, P, P, D, P, P, P, D, I,
                                           C.POINTER(C.POINTER(C.c_int)),
                                           C.POINTER(C.POINTER(C.c_int)),
                                           C.POINTER(C.POINTER(C.c_double)), P]
        L.orc_free.argtypes = [P]
        L.orc_fem_build_system_damped.argtypes = [I, I, P, P, P, P, P, P, D, P, P, P, P, D, I, D, D,
                                                  C.POINTER(C.POINTER(C.c_int)),
                                                  C.POINTER(C.POINTER(C.c_int)),
                                                  C.POINTER(C.POINTER(C.c_double)), P]
        L.orc_implicit_euler_velocity.argtypes = [I, P, P, P, D, P]
        L.orc_plane_collide.argtypes = [I, P, I, P, P, P, P]
        L.orc_plane_collide.restype = I
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class _Csr:
    """Keeps the numpy buffers of a CSR alive while C holds pointers."""

    def __init__(self, M, cols=None):
        M = sp.csr_matrix(M)
        M.sort_indices()
        self.ptr = _i32(M.indptr)
        self.idx = _i32(M.indices)
        self.val = _f64(M.data)
        self.s = _CSR(M.shape[0], M.shape[1] if cols is None else cols,
                      _ptr(self.ptr), _ptr(self.idx), _ptr(self.val))


def spmv(A, x):
    a = _Csr(A)
    x = _f64(x)
    y = np.empty(A.shape[0])
    lib().orc_spmv(C.byref(a.s), _ptr(x), _ptr(y))
    return y


def func_value(A, x, b):
    a = _Csr(A)
    x, b = _f64(x), _f64(b)
    return lib().orc_func_value(C.byref(a.s), _ptr(x), _ptr(b))


class Projector:
    """DecoupledConProjector restated (MPRGPProjection.h:320-397)."""

    def __init__(self, J, c):
        self._J = _Csr(J)
        self.c = _f64(c)
        self.n = J.shape[1]
        self.m = J.shape[0]
        self.h = lib().orc_projector_new(C.byref(self._J.s), _ptr(self.c))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_projector_free(self.h)
            self.h = None

    def is_decoupled(self):
        return bool(lib().orc_projector_is_decoupled(self.h))

    def project(self, x):
        x = _f64(x)
        y = np.empty(self.n)
        lib().orc_project(self.h, _ptr(x), _ptr(y))
        return y

    def decide_face(self, x):
        x = _f64(x)
        lib().orc_decide_face(self.h, _ptr(x))
        return self.face()

    def face(self):
        f = lib().orc_face(self.h)
        return np.array([f[i][0] if isinstance(f[i], bytes) else f[i] for i in range(self.m)],
                        dtype=np.int8) if self.m else np.zeros(0, np.int8)

    def set_face(self, face):
        f = np.ascontiguousarray(face, dtype=np.int8)
        lib().orc_set_face(self.h, _ptr(f))

    def num_active(self):
        return lib().orc_num_active(self.h)

    def is_feasible(self, x):
        x = _f64(x)
        return bool(lib().orc_is_feasible(self.h, _ptr(x)))

    def phi(self, g):
        g = _f64(g)
        out = np.empty(self.n)
        lib().orc_phi(self.h, _ptr(g), _ptr(out))
        return out

    def beta(self, g, phi):
        g, phi = _f64(g), _f64(phi)
        out = np.empty(self.n)
        lib().orc_beta(self.h, _ptr(g), _ptr(phi), _ptr(out))
        return out

    def phitphi(self, x, alpha_bar, phi):
        x, phi = _f64(x), _f64(phi)
        return lib().orc_phitphi(self.h, _ptr(x), float(alpha_bar), _ptr(phi))

    def step_limit(self, x, d, alpha_cg=np.finfo(np.float64).max):
        x, d = _f64(x), _f64(d)
        return lib().orc_step_limit(self.h, _ptr(x), _ptr(d), float(alpha_cg))


class Precond:
    """DiagonalDecouplePreconSolver restated (MPRGPPrecondition.h:72-110)."""

    def __init__(self, A, projector: Projector, preconditioned=True):
        self._A = _Csr(A)
        self.P = projector
        self.n = A.shape[0]
        self.h = lib().orc_precond_new(C.byref(self._A.s), projector.h, int(preconditioned))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_precond_free(self.h)
            self.h = None

    def solve(self, g, phi):
        g, phi = _f64(g), _f64(phi)
        z = np.empty(self.n)
        lib().orc_precond_solve(self.h, self.P.h, _ptr(g), _ptr(z), _ptr(phi))
        return z


def spec_rad(A, eps, ev=None, seed=0x5EED):
    """MPRGPBase::specRad restated; returns (lambda, iterations, ev_out or None)."""
    a = _Csr(A)
    n = A.shape[0]
    evb = np.zeros(n) if ev is None else _f64(ev).copy()
    valid = C.c_int(0 if ev is None else 1)
    it = C.c_int(0)
    lam = lib().orc_spec_rad(C.byref(a.s), _ptr(evb), C.byref(valid), float(eps),
                             C.c_ulonglong(seed), C.byref(it))
    return lam, it.value, (evb if valid.value else None)


def decoupled_solve(A, B, J, c, x, tol=1e-3, max_it=1000, variant=MONOTONIC,
                    preconditioned=True, alpha_bar=0.0, ev=None, project_first=True):
    """MPRGPDecoupledCon<double>::solve restated.  Returns (code, x, stats-dict, ev)."""
    a = _Csr(A)
    j = _Csr(J, cols=A.shape[0])
    B = _f64(B)
    c = _f64(c)
    x = _f64(x).copy()
    n = A.shape[0]
    evb = np.zeros(n) if ev is None else _f64(ev).copy()
    valid = C.c_int(0 if ev is None else 1)
    st = Stats()
    code = lib().orc_decoupled_solve(C.byref(a.s), _ptr(B), C.byref(j.s), _ptr(c), _ptr(x),
                                     float(tol), int(max_it), int(variant), int(preconditioned),
                                     float(alpha_bar), _ptr(evb), C.byref(valid),
                                     int(project_first), C.byref(st))
    return code, x, st.as_dict(), (evb if valid.value else None)


EV_RESIDUAL, EV_PROP, EV_CG_TEST, EV_FEASIBLE, EV_STEP, EV_Y_INFEASIBLE, EV_FY_CLOSED = range(7)
STEP_CG, STEP_EXP, STEP_PROP = range(3)
F_DIRECT, F_CLOSED, F_CLOSED_FROM_FPY = range(3)
EVENT_DTYPE = np.dtype([("iteration", np.int32), ("kind", np.int32), ("outcome", np.int32),
                        ("pad", np.int32), ("lhs", np.float64), ("rhs", np.float64)])


class _Options(C.Structure):
    _fields_ = [("events", C.c_void_p), ("cap", C.c_int), ("n", C.c_int), ("fmode", C.c_int)]


def decoupled_solve_traced(A, B, J, c, x, tol=1e-3, max_it=1000, preconditioned=True, alpha_bar=0.0,
                           ev=None, project_first=True, fmode=F_DIRECT, cap=1 << 20):
    """The monotone solve with its decisions recorded (orc_decoupled_solve_ex).
    Returns (code, x, stats-dict, events) with events a structured array of
    EVENT_DTYPE in the order the control flow took them."""
    a = _Csr(A)
    j = _Csr(J, cols=A.shape[0])
    B = _f64(B)
    c = _f64(c)
    x = _f64(x).copy()
    n = A.shape[0]
    evb = np.zeros(n) if ev is None else _f64(ev).copy()
    valid = C.c_int(0 if ev is None else 1)
    st = Stats()
    events = np.zeros(cap, EVENT_DTYPE)
    opt = _Options(events.ctypes.data, cap, 0, int(fmode))
    code = lib().orc_decoupled_solve_ex(C.byref(a.s), _ptr(B), C.byref(j.s), _ptr(c), _ptr(x),
                                        float(tol), int(max_it), MONOTONIC, int(preconditioned),
                                        float(alpha_bar), _ptr(evb), C.byref(valid),
                                        int(project_first), C.byref(opt), C.byref(st))
    if opt.n > cap:
        raise RuntimeError(f"{opt.n} events, capacity {cap}")
    return code, x, st.as_dict(), events[:opt.n].copy()


def step_sequence(events):
    """The step types (STEP_*) of a traced solve, in order."""
    return events["outcome"][events["kind"] == EV_STEP].copy()


# ---------------------------------------------------------- coupled rows

def qp_nonneg(G, h):
    """argmin 1/2 l'Gl - h'l s.t. l >= 0 (the reference's solveQP with J = I)."""
    G = np.ascontiguousarray(G, np.float64)
    h = _f64(h)
    lam = np.zeros(len(h))
    rc = lib().orc_qp_nonneg(len(h), _ptr(G), _ptr(h), _ptr(lam))
    return rc, lam


class GeneralProjector:
    """GeneralConProjector (MPRGPProjection.h:400-520) over rows J x >= c."""

    def __init__(self, J, c):
        self._j = _Csr(J)
        self.n = J.shape[1]
        self.m = J.shape[0]
        self._c = _f64(c).copy()
        self.h = lib().orc_gprojector_new(C.byref(self._j.s), _ptr(self._c))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_gprojector_free(self.h)
            self.h = None

    def project(self, x):
        y = np.zeros(self.n)
        rc = lib().orc_gproject(self.h, _ptr(_f64(x)), _ptr(y))
        return y if rc == 0 else None

    def decide_face(self, x):
        return lib().orc_gdecide_face(self.h, _ptr(_f64(x)))

    def face(self):
        f = lib().orc_gface(self.h)
        return np.array([f[i][0] for i in range(self.m)], dtype=np.int8) if self.m else np.zeros(0, np.int8)

    def num_active(self):
        return lib().orc_gnum_active(self.h)

    def is_feasible(self, x):
        return bool(lib().orc_gis_feasible(self.h, _ptr(_f64(x))))

    def phi(self, g):
        out = np.zeros(self.n)
        lib().orc_gphi(self.h, _ptr(_f64(g)), _ptr(out))
        return out

    def beta(self, g, phi):
        out = np.zeros(self.n)
        lib().orc_gbeta(self.h, _ptr(_f64(g)), _ptr(_f64(phi)), _ptr(out))
        return out

    def phitphi(self, x, alpha_bar, phi):
        return lib().orc_gphitphi(self.h, _ptr(_f64(x)), float(alpha_bar), _ptr(_f64(phi)))


def general_solve(A, B, J, c, x, tol=1e-3, max_it=1000, preconditioned=True, alpha_bar=0.0, ev=None,
                  project_first=True):
    """MPRGPGeneralCon<double>::solve restated.  Returns (code, x, stats-dict)."""
    a = _Csr(A)
    j = _Csr(J, cols=A.shape[0])
    B = _f64(B)
    c = _f64(c)
    x = _f64(x).copy()
    n = A.shape[0]
    evb = np.zeros(n) if ev is None else _f64(ev).copy()
    valid = C.c_int(0 if ev is None else 1)
    st = Stats()
    code = lib().orc_general_solve(C.byref(a.s), _ptr(B), C.byref(j.s), _ptr(c), _ptr(x), float(tol),
                                   int(max_it), int(preconditioned), float(alpha_bar), _ptr(evb),
                                   C.byref(valid), int(project_first), C.byref(st))
    return code, x, st.as_dict()


# ------------------------------------------------------------------ ICA

class IcaStats(C.Structure):
    _fields_ = [("code", C.c_int), ("iterations", C.c_int), ("gs_iterations", C.c_int),
                ("residual", C.c_double), ("gs_residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def inv3(m):
    """3x3 inverse by the adjugate (the ICA comparator's Da^-1 blocks)."""
    m = _f64(np.asarray(m).reshape(9))
    out = np.empty(9)
    lib().orc_inv3(_ptr(m), _ptr(out))
    return out.reshape(3, 3)


def pgs(B, c, lam0=None, max_it=20, tol=1e-3):
    """ProjectedGaussSeidel::solve (ICASolver.h:68-111) on B (CSR) from lam0.
    Returns (tolerance met, lambda, sweep index at the stop, residual)."""
    b = _Csr(B)
    c = _f64(c)
    lam = np.zeros(B.shape[0]) if lam0 is None else _f64(lam0).copy()
    it, res = C.c_int(), C.c_double()
    met = lib().orc_pgs(C.byref(b.s), _ptr(c), _ptr(lam), int(max_it), float(tol), C.byref(it), C.byref(res))
    return bool(met), lam, it.value, res.value


def ica_b(J, n_nodes, dinv):
    """B = J Da^-1 J^T (scipy CSR) from the per-node inverse blocks dinv (n_nodes x 3 x 3)."""
    j = _Csr(J, cols=3 * n_nodes)
    d = _f64(np.asarray(dinv).reshape(-1))
    nnz = C.c_int()
    pp, pi, pv = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    lib().orc_ica_b(C.byref(j.s), int(n_nodes), _ptr(d), C.byref(nnz), C.byref(pp), C.byref(pi), C.byref(pv))
    m = J.shape[0]
    ptr = np.ctypeslib.as_array(pp, (m + 1,)).copy()
    idx = np.ctypeslib.as_array(pi, (max(nnz.value, 1),))[:nnz.value].copy()
    val = np.ctypeslib.as_array(pv, (max(nnz.value, 1),))[:nnz.value].copy()
    for q in (pp, pi, pv):
        lib().orc_ica_free(C.cast(q, C.c_void_p))
    return sp.csr_matrix((val, idx, ptr), shape=(m, m))


def ica_solve(A, R, J, p, x, tol=1e-4, max_it=1000, gs_max_it=20, gs_tol=1e-3):
    """ICASolver(max_it, tol, gs_max_it, gs_tol).reset(A, R) + solve(J, p, x)
    (ICASolver.h:287-415) restated.  Returns (code, x, lambda, stats-dict)."""
    a = _Csr(A)
    n = A.shape[0]
    J = sp.csr_matrix((0, n)) if J is None else J
    j = _Csr(J, cols=n)
    R, p = _f64(R), _f64(p) if len(p) else np.zeros(1)
    x = _f64(x).copy()
    lam = np.zeros(max(J.shape[0], 1))
    st = IcaStats()
    code = lib().orc_ica_solve(C.byref(a.s), _ptr(R), C.byref(j.s), _ptr(p), _ptr(x), float(tol), int(max_it),
                               int(gs_max_it), float(gs_tol), _ptr(lam), C.byref(st))
    return code, x, lam[:J.shape[0]], st.as_dict()


# ------------------------------------------------------------------ FEM

def fem_setup(n_nodes, tets, X0):
    tets = _i32(tets).copy()
    X0 = _f64(X0)
    T = tets.shape[0]
    d = np.empty((T, 9))
    vol = np.empty(T)
    mass = np.empty(n_nodes)
    lib().orc_fem_setup(int(n_nodes), int(T), _ptr(tets), _ptr(X0), _ptr(d), _ptr(vol), _ptr(mass))
    return tets, d, vol, mass


def polar_rotation(F):
    F = _f64(F).reshape(9)
    R = np.empty(9)
    lib().orc_polar_rotation(_ptr(F), _ptr(R))
    return R.reshape(3, 3)


def fem_element(mtype, xe, d, vol, lam, mu):
    xe = _f64(xe).reshape(12)
    d = _f64(d).reshape(9)
    fB = np.empty(12)
    HB = np.empty(144)
    E = lib().orc_fem_element(int(mtype), _ptr(xe), _ptr(d), float(vol), float(lam), float(mu),
                              _ptr(fB), _ptr(HB))
    return E, fB, HB.reshape(12, 12)


def fem_build_system(n_nodes, tets_oriented, d, vol, mass, young, poisson, density,
                     x_n, v_n, gravity, h, mtype, x1=None, alpha_k=0.0, alpha_m=0.0):
    """Returns (A as scipy CSR, b) at the Newton iterate x1 (None: x_n), with
    Rayleigh damping alpha_k (mass) and alpha_m (stiffness)."""
    tets = _i32(tets_oriented)
    T = tets.shape[0]
    d, vol, mass = _f64(d), _f64(vol), _f64(mass)
    young = _f64(np.broadcast_to(young, (T,)))
    poisson = _f64(np.broadcast_to(poisson, (T,)))
    x_n, v_n, gravity = _f64(x_n), _f64(v_n), _f64(gravity)
    x1 = x_n if x1 is None else _f64(x1)
    n = 3 * n_nodes
    b = np.empty(n)
    pp = C.POINTER(C.c_int)()
    ip = C.POINTER(C.c_int)()
    vp = C.POINTER(C.c_double)()
    L = lib()
    L.orc_fem_build_system_damped(int(n_nodes), int(T), _ptr(tets), _ptr(d), _ptr(vol), _ptr(mass),
                                  _ptr(young), _ptr(poisson), float(density), _ptr(x_n), _ptr(v_n),
                                  _ptr(x1), _ptr(gravity), float(h), int(mtype), float(alpha_k),
                                  float(alpha_m), C.byref(pp), C.byref(ip), C.byref(vp), _ptr(b))
    ptr = np.ctypeslib.as_array(pp, shape=(n + 1,)).copy()
    nnz = int(ptr[-1])
    idx = np.ctypeslib.as_array(ip, shape=(nnz,)).copy()
    val = np.ctypeslib.as_array(vp, shape=(nnz,)).copy()
    L.orc_free(C.cast(pp, C.c_void_p))
    L.orc_free(C.cast(ip, C.c_void_p))
    L.orc_free(C.cast(vp, C.c_void_p))
    return sp.csr_matrix((val, idx, ptr), shape=(n, n)), b


def plane_collide(x, planes):
    """Rows of the decoupled plane collision at positions x (N, 3).

    planes: (K, 5) rows {nx, ny, nz, p, delta}.  Returns (node, normal (m, 3),
    c) in the reference's push order, or raises ValueError on a non-unit
    normal (the reference's assertion)."""
    x = _f64(np.reshape(x, (-1, 3)))
    N = x.shape[0]
    planes = _f64(np.reshape(planes, (-1, 5)))
    node = np.zeros(max(N, 1), np.int32)
    normal = np.zeros((max(N, 1), 3))
    c = np.zeros(max(N, 1))
    m = lib().orc_plane_collide(N, _ptr(x), planes.shape[0], _ptr(planes), _ptr(node), _ptr(normal), _ptr(c))
    if m < 0:
        raise ValueError("plane normal is not unit length")
    return node[:m].copy(), normal[:m].copy(), c[:m].copy()


def implicit_euler_velocity(x_n, v_n, x, h):
    """updateMesh's velocity under implicit Euler (MprgpFemSolver.cpp:377-390)."""
    x_n, v_n, x = _f64(x_n), _f64(v_n), _f64(x)
    v = np.empty(x.size)
    lib().orc_implicit_euler_velocity(int(x.size), _ptr(x_n), _ptr(v_n), _ptr(x), float(h), _ptr(v))
    return v


def advance(n_nodes, tets_oriented, d, vol, mass, young, poisson, density, x_n, v_n, gravity, h, mtype,
            planes, newton_max_it=10, newton_eps=1e-4, tol=1e-6, max_it=1000, alpha_bar=0.0, ev=None,
            alpha_k=0.0, alpha_m=0.0):
    """One time step of MprgpFemSolver::advance (MprgpFemSolver.cpp:193-203)
    with the simulator's implicit Euler (Simulator.cpp:313): the planes'
    collision at x_n (DCDhandle, :255-271), DecoupledMprgpFemSolver::forward's
    Newton loop (:514-537: buildLinearSystem at x1, x = P(x1), MPRGP solve,
    updatePos :358-375 until max|x - x1| < eps or maxIter solves), then
    updateMesh (:377-390).  Returns (x, v, info dict)."""
    x_n, v_n = _f64(x_n).ravel(), _f64(v_n).ravel()
    node, normal, c = plane_collide(x_n.reshape(-1, 3), planes)
    m = node.size
    J = sp.csr_matrix((normal.ravel(), (np.repeat(np.arange(m), 3),
                                        (3 * node[:, None] + np.arange(3)).ravel())),
                      shape=(m, 3 * n_nodes))
    x1 = x_n.copy()
    info = {"rows": m, "newton_iters": 0, "codes": [], "iterations": 0, "delta": 0.0}
    for it in range(newton_max_it):
        A, b = fem_build_system(n_nodes, tets_oriented, d, vol, mass, young, poisson, density, x_n, v_n,
                                gravity, h, mtype, x1=x1, alpha_k=alpha_k, alpha_m=alpha_m)
        code, x, st, ev = decoupled_solve(A, b, J, c, x1, tol=tol, max_it=max_it, alpha_bar=alpha_bar, ev=ev,
                                          project_first=True)
        info["codes"].append(code)
        info["iterations"] += st["iterations"]
        info["newton_iters"] = it + 1
        delta = float(np.max(np.abs(x - x1))) if x.size else 0.0
        info["delta"] = delta
        x1 = x
        if delta < newton_eps:
            break
    info["ev"] = ev
    return x1, implicit_euler_velocity(x_n, v_n, x1, h), info
