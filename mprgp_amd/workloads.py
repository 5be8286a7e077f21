"""Device-resident QPs built from synthetic scenes, and the algorithmic-byte
accounting used for roofline figures.

``device_qp`` assembles a scene's system on the GPU (``Fem.assemble``), sets
the collider's rows and the warm start ``x_n``; the solve projects it first
(``project_first``), as DecoupledMprgpFemSolver::forward does before calling
MPRGPDecoupledCon::solve (MprgpFemSolver.cpp:528-533).
"""
from __future__ import annotations

import numpy as np

import mprgp_amd as M
from mprgp_amd import problems

# op indices of the device state machine (mprgp_dev.h)
OPS = ["DONE", "PROJX", "POW_IT", "POW_SAVE", "INIT", "TOP", "CG_AP", "CG_PY", "CG_Y",
       "CG_A", "EXP1", "EXP2", "EXP3", "PROP1", "PROP2", "FINAL", "CL_AP", "CL_CG",
       "CL_EXP", "CL_EXP2", "CL_PROP1", "R_START", "R_UPD", "R_FINAL", "CG_PYG",
       "CG_BAP", "CG_YA"]


class DeviceQP:
    def __init__(self, ctx, scene: problems.Scene, material=None):
        self.scene = scene
        self.fem = M.Fem(ctx, scene.X0, scene.tets, scene.young, scene.poisson, scene.density)
        self.qp = self.fem.matrix.qp()
        self.fem.assemble(self.qp, scene.x_n.ravel(), scene.v_n.ravel(), scene.gravity, scene.h,
                          scene.material if material is None else material)
        node, normal, c = scene.rows()
        self.m = int(node.size)
        self.qp.set_rows(node, normal, c)
        self.qp.set_x(scene.x_n.ravel())
        info = self.fem.matrix.info()
        self.n_nodes = info["n_nodes"]
        self.n_blocks = info["n_blocks"]
        self.alpha_bar = 0.0


# Streams each op touches, every stream counted once (ideal traffic): "M" the
# matrix of the op's precision form and "M64" the FP64 matrix (values, column
# indices of `idx` bytes, slice offsets; a fused two-vector SpMV still reads
# it once);
# vector names stand for 3 values per node (gathered SpMV inputs count once);
# "rows" is the per-node flag byte plus the 6 FP64 words of each constraint
# row (normal, offset, |n|^2, tangent-Jacobi factor); "cc" one FP64 offset per
# row; "face" one byte per node.
_STREAMS = {
    "PROJX": ("rows", "x", "x"),
    "POW_IT": ("M", "w", "v", "v", "w"),
    "POW_SAVE": ("pv", "ev"),
    "INIT": ("M", "x", "b", "rows", "dinv", "g", "phi", "z", "p", "face", "beta"),
    "EXP3": ("M", "x", "b", "rows", "dinv", "g", "phi", "z", "p", "face", "py", "y", "beta"),
    "CL_EXP2": ("M", "x", "b", "rows", "dinv", "g", "phi", "z", "p", "face"),
    "TOP": ("rows", "face", "g", "x", "phi", "beta"),
    "CG_AP": ("M", "p", "ap", "z", "g"),
    "CL_AP": ("M", "p", "ap", "z", "g", "rows", "x"),
    "CG_PY": ("rows", "x", "p", "y", "py"),
    "CG_Y": ("M", "py", "b"),
    "CG_A": ("rows", "face", "ap", "g", "dinv", "x", "g", "phi", "z", "beta"),
    "CG_PYG": ("rows", "face", "x", "p", "y", "g", "ap", "dinv", "g", "phi", "z", "beta"),
    "CG_BAP": ("M", "z", "p", "p", "ap"),
    "CG_YA": ("M", "py", "b", "rows", "face", "ap", "g", "dinv", "y", "g", "phi", "z", "beta"),
    "EXP1": ("rows", "x", "y"),
    "EXP2": ("M", "y", "b", "rows", "g", "py", "x", "face"),
    "PROP1": ("M", "x", "beta", "b", "g", "ap"),
    "PROP2": ("rows", "x", "beta", "g", "ap", "dinv", "x", "g", "phi", "z", "p", "face", "beta"),
    "FINAL": ("rows", "x", "face", "x"),
    "CL_CG": ("rows", "face", "ap", "x", "p", "g", "dinv", "x", "g", "phi", "z"),
    "CL_EXP": ("rows", "x", "p", "g", "ap", "x", "g", "face"),
    "CL_PROP1": ("M", "beta", "g", "ap"),
    "R_START": ("M64", "x64", "b64", "rows", "face", "b", "x", "cc"),
    "R_UPD": ("x64", "x", "x64"),
    "R_FINAL": ("rows", "x64", "face", "x64"),
}

# FP32 streams of the mixed form's inner solve; everything else (x, y, P(y),
# 1/diag, the rows) is FP64.
MIXED_F32 = {"M", "b", "g", "phi", "beta", "p", "z", "ap"}
# Not counted (a lower bound): CG_PYG's matrix and neighbour reads on the few
# rows the projection moves, and CG_AP's reads of x and b at CG entry (the
# mixed form's f(x)).


def index_bytes(n_nodes, slices_per_wave=1):
    """Bytes of one column index in the algorithmic count: 16-bit for
    matrices of at most 65536 nodes at one slice per wavefront (the step
    kernel reads those when every QP of a launch qualifies; its tail form, for
    launches of few tiles, reads the 32-bit copy, and those 2 extra bytes
    count as traffic, not as algorithmic bytes), else 32-bit."""
    return 2 if n_nodes <= 65536 and slices_per_wave == 1 else 4


def spmv_bytes(N, nb, s=8, idx=4):
    """Matrix once, input vector once, output vector once."""
    return nb * (9 * s + idx) + 4 * (N + 1) + 2 * 3 * N * s


def op_bytes(op, N, nb, m, mixed=False, precond=True, idx=4):
    """Algorithmic HBM bytes of one device op on one QP (ideal traffic: every
    stream read or written once).  ``mixed``: the inner-solve ops read the
    FP32 matrix and FP32 direction vectors; ``idx``: bytes per column index."""
    name = OPS[op]
    total = 0
    streams = _STREAMS.get(name, ())
    for st in streams:
        if st == "rows":
            total += N + 48 * m
        elif st == "cc":
            total += 8 * m
        elif st == "face":
            total += N
        elif st in ("M", "M64"):
            if mixed and name == "INIT":
                continue  # the correction QP starts from d = 0: no product with A
            s = 4 if (mixed and st == "M") else 8
            total += nb * (9 * s + idx) + 4 * (N + 1)
        elif st == "dinv" and not precond:
            continue
        else:
            s = 4 if (mixed and st in MIXED_F32) else 8
            total += 3 * N * s
    return total


def solve_bytes(counts, N, nb, m, mixed=False, precond=True, idx=4):
    """Algorithmic bytes of a whole solve from its per-op launch counts."""
    total = 0
    for op, k in enumerate(counts[:len(OPS)]):
        if k:
            total += int(k) * op_bytes(op, N, nb, m, mixed, precond, idx)
    return total
