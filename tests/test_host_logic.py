"""Host-side logic that needs no GPU: the op table shared by the device state
machine and the byte accounting, the accounting itself, the serving
workload's scenes, and the autotuner's cache keys."""
import os
import re

import numpy as np
import pytest

import mprgp_amd as M
from mprgp_amd import problems, serving, tune, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def device_ops():
    text = open(os.path.join(ROOT, "mprgp_amd", "csrc", "mprgp_dev.h")).read()
    body = text[text.index("enum Op : int {"):text.index("OP_NUM")]
    return re.findall(r"\bOP_([A-Z0-9_]+)\b", body)


def test_op_names_match_device_enum():
    assert device_ops() == workloads.OPS


def test_byte_accounting():
    N, nb, m = 10000, 150000, 500
    mat64 = nb * (9 * 8 + 4) + 4 * (N + 1)
    mat32 = nb * (9 * 4 + 4) + 4 * (N + 1)
    v64, v32 = 24 * N, 12 * N
    op = workloads.OPS.index
    # a power iteration: the form's matrix, w_{k-1} (gathered) and v_{k-1}
    # read, v_k and w_k written, all FP64
    assert workloads.op_bytes(op("POW_IT"), N, nb, m) == mat64 + 4 * v64
    assert workloads.op_bytes(op("POW_IT"), N, nb, m, True) == mat32 + 4 * v64
    # CG_AP: matrix + p (gathered once) + AP + z + g
    assert workloads.op_bytes(op("CG_AP"), N, nb, m) == mat64 + 4 * v64
    assert workloads.op_bytes(op("CG_AP"), N, nb, m, True) == mat32 + 4 * v32  # all FP32 in the inner solve
    # the mixed form's fused CG step: rows (flag byte + 6 words per row), face,
    # x, y and 1/diag in FP64; p, g read and written, AP, phi, z, beta in FP32
    assert workloads.op_bytes(op("CG_PYG"), N, nb, m, True) == (N + 48 * m) + N + 3 * v64 + 7 * v32
    # no preconditioner: no D^-1 stream
    assert (workloads.op_bytes(op("CG_A"), N, nb, m) - workloads.op_bytes(op("CG_A"), N, nb, m, precond=False)
            == v64)
    # the mixed form's INIT starts from d = 0 and reads no matrix
    assert workloads.op_bytes(op("INIT"), N, nb, m) - workloads.op_bytes(op("INIT"), N, nb, m, True) > mat32
    # 16-bit column indices: 2 bytes less per block
    assert workloads.op_bytes(op("CG_AP"), N, nb, m, True, idx=2) == mat32 - 2 * nb + 4 * v32
    assert workloads.index_bytes(65536) == 2 and workloads.index_bytes(65537) == 4
    assert workloads.index_bytes(100, slices_per_wave=2) == 4
    counts = np.zeros(32, np.int32)
    counts[op("CG_AP")] = 3
    counts[op("TOP")] = 2
    assert workloads.solve_bytes(counts, N, nb, m) == 3 * workloads.op_bytes(op("CG_AP"), N, nb, m) + \
        2 * workloads.op_bytes(op("TOP"), N, nb, m)


def test_serving_scenes_start_clear_of_their_planes():
    rng = np.random.default_rng(5)
    for i in range(12):
        sc = serving.session_scene(rng, i, dof_range=(3e3, 2e4))
        assert sc.material == M.COROTATIONAL
        assert np.all(problems.signed_volumes(sc.X0, sc.tets) > 0)
        for n, p, delta in sc.planes:
            n = np.asarray(n) / np.linalg.norm(n)
            s = sc.x_n @ n + p
            assert s.min() > 0.0            # no node starts inside a plane
            assert s.min() <= delta + 0.02  # ... and the body is close enough to land on it


def test_tune_keys_and_shapes():
    class Q:
        def __init__(self, n):
            self.n = n
    a = [Q(3 * 1000)] * 10
    b = [Q(3 * 1000)] * 10 + [Q(30)]
    assert tune.batch_key(a, M.MIXED) == tune.batch_key(a, M.MIXED)
    assert tune.batch_key(a, M.MIXED) != tune.batch_key(a, M.FP64)
    assert tune.batch_key(a, M.MIXED).startswith(f"v{tune.CACHE_VERSION}/dev0/mixed/")
    assert tune.batch_key(b, M.MIXED) == tune.batch_key(a, M.MIXED)  # same size class
    p = tune.apply(M.params(), {"tiles": 8, "graph_steps": 32})
    assert (p.tiles, p.graph_steps, p.tail_tiles) == (8, 32, 0)
    p = tune.apply(M.params(), {"tiles": 2, "graph_steps": 16, "tail_tiles": 1024, "xcd_layout": 1,
                                "slices_per_wave": 2})
    assert (p.tiles, p.graph_steps, p.tail_tiles, p.xcd_layout, p.slices_per_wave) == (2, 16, 1024, 1, 2)
    assert tune.apply(M.params(), {"tiles": 2, "graph_steps": 16}).slices_per_wave == 1
    assert set(tune.SPW_CANDIDATES) <= {1, 2}
    assert 0 in tune.TAIL_CANDIDATES and all(t >= 0 for t in tune.TAIL_CANDIDATES)
    for s in tune.DEFAULT_CANDIDATES:
        assert s["tiles"] in (1, 2, 4, 8) and s["graph_steps"] > 0


def test_batch_stats_sequence():
    """solve_batch's stats: one dict per QP, made from the C array on first
    access; indexing, negative indices, slices, iteration and len behave as
    for a list."""
    import mprgp_amd as M
    raw = (M.Stats * 3)()
    for i in range(3):
        raw[i].code, raw[i].iterations, raw[i].residual = -i, 10 * i, 0.5 * i
    st = M.BatchStats(raw)
    assert len(st) == 3
    assert st[1]["iterations"] == 10 and st[-1]["code"] == -2 and st[1] is st[1]
    assert [s["residual"] for s in st] == [0.0, 0.5, 1.0]
    assert [s["code"] for s in st[1:]] == [-1, -2]
    assert sum(1 for s in st if s["code"] == 0) == 1
    assert "reserved_d" not in st[0] and set(st[0]) == set(raw[0].as_dict())
    with pytest.raises(IndexError):
        st[3]
    with pytest.raises(IndexError):
        st[-4]
