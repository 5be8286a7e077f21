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
OPS = ["DONE", "PROJX", "POW_MV", "POW_NRM", "POW_SAVE", "INIT", "TOP", "CG_AP", "CG_PY", "CG_Y",
       "CG_A", "CG_B", "EXP1", "EXP2", "EXP3", "PROP1", "PROP2", "FINAL", "CL_AP", "CL_CG",
       "CL_EXP", "CL_EXP2", "CL_PROP1"]


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


def spmv_bytes(N, nb, s=8):
    """Matrix once, input vector once, output vector once (index bytes int32)."""
    return nb * (9 * s + 4) + 4 * (N + 1) + 2 * 3 * N * s


def op_bytes(op, N, nb, m, s=8, precond=True):
    """Algorithmic HBM bytes of one device op on one QP (ideal traffic: every
    stream read or written once)."""
    v = 3 * N * s
    rows = N + m * 6 * s          # flag per node, normal/c/|n|^2/factor per row
    face = N
    dinv = v if precond else 0
    mat = nb * (9 * s + 4) + 4 * (N + 1)
    spmv = spmv_bytes(N, nb, s)
    name = OPS[op]
    table = {
        "PROJX": 2 * v + rows,
        "POW_MV": spmv,
        "POW_NRM": 3 * v,
        "POW_SAVE": 2 * v,
        "INIT": spmv + v + 3 * v + rows + face + dinv,
        "EXP3": spmv + v + 3 * v + rows + face + dinv + 2 * v,
        "CL_EXP2": spmv + v + 3 * v + rows + face + dinv,
        "TOP": 5 * v + face + rows,
        "CG_AP": spmv + 2 * v,
        "CL_AP": spmv + 3 * v + rows,
        "CG_PY": 4 * v + rows,
        "CG_Y": mat + v + v,
        "CG_A": 7 * v + dinv + face + rows,
        "CG_B": 7 * v + face + rows,
        "EXP1": 2 * v + rows,
        "EXP2": spmv + v + 2 * v + face + rows,
        "PROP1": mat + 2 * v + 2 * v + v,
        "PROP2": 9 * v + face + rows + dinv,
        "FINAL": 2 * v + rows + face,
        "CL_CG": 8 * v + dinv + face + rows,
        "CL_EXP": 6 * v + face + rows,
        "CL_PROP1": spmv + v,
    }
    return table.get(name, 0)


def solve_bytes(counts, N, nb, m, s=8, precond=True, cg_entries=0):
    """Algorithmic bytes of a whole solve from its per-op launch counts."""
    total = 0
    for op, k in enumerate(counts[:len(OPS)]):
        if k:
            total += int(k) * op_bytes(op, N, nb, m, s, precond)
    # a CG entry evaluates f(x) as well: a second input vector on CG_Y
    total += cg_entries * 3 * N * s
    return total
