"""ctypes bindings of ``libmprgp_amd.so`` (the C ABI in ``include/mprgp_amd.h``).

The library is built in-tree by ``mprgp_amd/build.py``; loading fails loudly
when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MPRGP_LIB") or os.path.join(HERE, "libmprgp_amd.so")

OK = 0
NOT_CONVERGED = -1
EBADINPUT = -2
ENOMEM = -3
EHIP = -4
ENAN = -5
ENODEV = -6
EBUDGET = -7
ESTATE = -8
EPENDING = -9

MONOTONIC = 0
CLASSIC = 1
FP64 = 0
MIXED = 1
LINEAR = 1
COROTATIONAL = 2
RHS_ONLY = 0x100


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_it", C.c_int), ("variant", C.c_int),
                ("precision", C.c_int), ("preconditioned", C.c_int), ("project_first", C.c_int),
                ("use_ev", C.c_int), ("alpha_bar", C.c_double), ("power_tol", C.c_double),
                ("tiles", C.c_int), ("graph_steps", C.c_int), ("max_refine", C.c_int),
                ("tail_tiles", C.c_int), ("xcd_layout", C.c_int), ("slices_per_wave", C.c_int),
                ("from_x0", C.c_int), ("reserved_i", C.c_int * 1),
                ("inner_tol", C.c_double),
                ("reserved_d", C.c_double * 3)]


class Stats(C.Structure):
    _fields_ = [("code", C.c_int), ("iterations", C.c_int), ("num_cg", C.c_int),
                ("num_exp", C.c_int), ("num_prop", C.c_int), ("active", C.c_int),
                ("power_iters", C.c_int), ("ev_valid", C.c_int), ("refinements", C.c_int),
                ("steps", C.c_int), ("launches", C.c_int), ("fp64_finish", C.c_int), ("residual", C.c_double),
                ("alpha_bar", C.c_double), ("lambda_max", C.c_double),
                ("elapsed_ms", C.c_double), ("queue_ms", C.c_double), ("error_bound", C.c_double),
                ("reserved_d", C.c_double * 2)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class StepParams(C.Structure):
    _fields_ = [("h", C.c_double), ("gravity", C.c_double * 3), ("material", C.c_int),
                ("n_planes", C.c_int), ("planes", C.c_void_p), ("newton_max_it", C.c_int),
                ("newton_eps", C.c_double), ("reserved_i", C.c_int * 4), ("reserved_d", C.c_double * 4)]


class StepStats(C.Structure):
    _fields_ = [("newton_iters", C.c_int), ("rows", C.c_int), ("code", C.c_int), ("iterations", C.c_int),
                ("delta", C.c_double), ("last", Stats)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "last"}
        d["last"] = self.last.as_dict()
        return d


class EngineConfig(C.Structure):
    _fields_ = [("max_active", C.c_int), ("tiles", C.c_int), ("graph_steps", C.c_int),
                ("reserved_i", C.c_int * 5), ("hbm_budget", C.c_size_t), ("reserved_d", C.c_double * 4)]


class EngineInfo(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("rounds", C.c_longlong),
                ("launches", C.c_longlong), ("active", C.c_int), ("queued", C.c_int),
                ("peak_active", C.c_int), ("reserved_i", C.c_int * 5), ("active_bytes", C.c_size_t),
                ("peak_bytes", C.c_size_t), ("reserved_d", C.c_double * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class MprgpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mprgp error {code}: {msg}")
        self.code = code


_lib = None

P = C.c_void_p
I = C.c_int
D = C.c_double
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every entry point of include/mprgp_amd.h
SIGNATURES = {
    "mprgp_params_default": (None, [C.POINTER(Params)]),
    "mprgp_last_error": (C.c_char_p, []),
    "mprgp_device_count": (I, [C.POINTER(I)]),
    "mprgp_ctx_create": (I, [I, C.c_size_t, PP]),
    "mprgp_ctx_destroy": (None, [P]),
    "mprgp_ctx_memory": (I, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mprgp_matrix_create_bsr3": (I, [P, I, P, P, P, PP]),
    "mprgp_matrix_create_csr": (I, [P, I, P, P, P, PP]),
    "mprgp_matrix_update_bsr3": (I, [P, P]),
    "mprgp_matrix_info": (I, [P, C.POINTER(I), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "mprgp_matrix_set_lambda_min": (I, [P, C.c_double]),
    "mprgp_matrix_get_bsr3": (I, [P, P, P, P]),
    "mprgp_general_solve": (I, [P, I, P, P, P, P, C.POINTER(Params), C.POINTER(Stats)]),
    "mprgp_ica_solve": (I, [P, I, P, P, P, P, C.c_double, I, I, C.c_double, P, C.POINTER(I), C.POINTER(Stats)]),
    "mprgp_matrix_spmv": (I, [P, P, P]),
    "mprgp_matrix_destroy": (None, [P]),
    "mprgp_qp_create": (I, [P, PP]),
    "mprgp_qp_destroy": (None, [P]),
    "mprgp_qp_set_rows": (I, [P, I, P, P, P]),
    "mprgp_qp_set_rows_sparse": (I, [P, I, I, P, P, P, P]),
    "mprgp_qp_collide_planes": (I, [P, I, P, P]),
    "mprgp_qp_get_rows": (I, [P, I, P, P, P, P]),
    "mprgp_qp_project": (I, [P, P, P]),
    "mprgp_matrix_func_value": (I, [P, P, P, P]),
    "mprgp_matrix_diagonal": (I, [P, P]),
    "mprgp_qp_is_feasible": (I, [P, P, P]),
    "mprgp_qp_decide_face": (I, [P, P, P]),
    "mprgp_qp_phi": (I, [P, P, P]),
    "mprgp_qp_beta": (I, [P, P, P, P]),
    "mprgp_qp_phitphi": (I, [P, P, D, P, P]),
    "mprgp_qp_step_limit": (I, [P, P, P, D, P]),
    "mprgp_qp_precond": (I, [P, P, P, P]),
    "mprgp_qp_set_rhs": (I, [P, P]),
    "mprgp_qp_get_rhs": (I, [P, P]),
    "mprgp_qp_set_x": (I, [P, P]),
    "mprgp_qp_reset_x": (I, [P]),
    "mprgp_qp_get_x": (I, [P, P]),
    "mprgp_qp_set_ev": (I, [P, P]),
    "mprgp_qp_get_ev": (I, [P, P, C.POINTER(I)]),
    "mprgp_qp_get_face": (I, [P, P]),
    "mprgp_qp_op_counts": (I, [P, P, I]),
    "mprgp_qp_solve": (I, [P, C.POINTER(Params), C.POINTER(Stats)]),
    "mprgp_qp_solve_noalloc": (I, [P, C.POINTER(Params), C.POINTER(Stats)]),
    "mprgp_solve_batch": (I, [P, I, PP, C.POINTER(Params), I, C.POINTER(Stats)]),
    "mprgp_decoupled_solve": (I, [I, I, P, P, P, P, I, P, P, P, P, P, D, I, P, C.POINTER(I),
                                  C.POINTER(Stats)]),
    "mprgp_qp_file_info": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_longlong),
                               C.POINTER(C.c_longlong)]),
    "mprgp_qp_file_read": (I, [C.c_char_p, P, P, P, P, P, P, P, P, P]),
    "mprgp_write_qp": (I, [C.c_char_p, I, P, P, P, P, I, P, P, P, P, P]),
    "mprgp_decoupled_solve_file": (I, [I, C.c_char_p, P, I, C.POINTER(I), D, I, P, C.POINTER(I),
                                       C.POINTER(Stats)]),
    "mprgp_fem_create": (I, [P, I, I, P, P, P, P, D, PP]),
    "mprgp_fem_assemble": (I, [P, I, P, P, P, D, P]),
    "mprgp_fem_assemble_at": (I, [P, I, P, P, P, P, D, P]),
    "mprgp_fem_set_damping": (I, [P, D, D]),
    "mprgp_fem_set_state": (I, [P, P, P]),
    "mprgp_fem_get_state": (I, [P, P, P]),
    "mprgp_fem_advance": (I, [P, P, C.POINTER(StepParams), C.POINTER(Params), C.POINTER(StepStats)]),
    "mprgp_fem_begin_step": (I, [P, P, C.POINTER(StepParams), P]),
    "mprgp_fem_end_step": (I, [P, P, C.POINTER(StepParams), P]),
    "mprgp_fem_matrix": (P, [P]),
    "mprgp_fem_destroy": (None, [P]),
    "mprgp_copy_probe": (I, [P, C.c_size_t, I, C.POINTER(D)]),
    "mprgp_engine_config_default": (None, [C.POINTER(EngineConfig)]),
    "mprgp_engine_create": (I, [P, C.POINTER(EngineConfig), PP]),
    "mprgp_engine_submit": (I, [P, P, C.POINTER(Params), C.POINTER(C.c_longlong)]),
    "mprgp_engine_wait": (I, [P, C.c_longlong, I, C.POINTER(Stats)]),
    "mprgp_engine_wait_any": (I, [P, I, C.POINTER(C.c_longlong), C.POINTER(Stats)]),
    "mprgp_engine_get_info": (I, [P, C.POINTER(EngineInfo)]),
    "mprgp_engine_destroy": (None, [P]),
}


def load():
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run mprgp_amd/build.py (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    s = load().mprgp_last_error()
    return s.decode() if s else ""


def check(rc: int, allow=(OK,)):
    if rc not in allow:
        raise MprgpError(rc, last_error())
    return rc
