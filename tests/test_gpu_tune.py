"""The batched solve's launch-shape autotuner (mprgp_amd/tune.py) on a GPU:
the tail switch point changes no bits, the tuned shape is cached and found
again, and every candidate shape's solution lies within the solves' own
error bounds of the others."""
import json

import numpy as np
import pytest

import mprgp_amd as M
from mprgp_amd import problems, tune, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch(ctx):
    scenes = problems.batch_scenes(16, seed=37, tol=1e-6, max_it=4000)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=1e-6, max_it=0, power_tol=1e-6))
    return qps, [s["alpha_bar"] for s in st]


def per_qp(ab, prec, **kw):
    return [M.params(tol=1e-6, max_it=4000, alpha_bar=a, project_first=1, precision=prec, **kw) for a in ab]


@pytest.mark.parametrize("prec", [M.FP64, M.MIXED])
def test_tail_switch_point_and_placement_change_no_bits(ctx, batch, prec):
    """tail_tiles = 1 (the tail form almost never), 0 (the default, 256) and
    2^30 (every launch), and both XCD placements, give the same bits and
    iteration counts."""
    qps, ab = batch
    out = {}
    for key in ((1, 0), (0, 0), (1 << 30, 0), (0, 1), (1 << 30, 1)):
        tail, lay = key
        for q in qps:
            q.qp.reset_x()
        rc, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per_qp(ab, prec, tiles=2, graph_steps=16,
                                                                        tail_tiles=tail, xcd_layout=lay))
        assert all(s["code"] == 0 for s in st), key
        out[key] = ([q.qp.get_x() for q in qps], [s["iterations"] for s in st])
    for key in list(out)[1:]:
        for a, b in zip(out[(1, 0)][0], out[key][0]):
            np.testing.assert_array_equal(a, b)
        assert out[key][1] == out[(1, 0)][1]


def test_tuner_caches_and_candidates_agree(ctx, batch, tmp_path):
    qps, ab = batch
    path = str(tmp_path / "tune.json")
    cands = [{"tiles": 2, "graph_steps": 16}, {"tiles": 4, "graph_steps": 8}, {"tiles": 1, "graph_steps": 32}]
    base = per_qp(ab, M.FP64)
    saved = (tune.DEFAULT_CANDIDATES, tune.SPW_CANDIDATES, tune.TAIL_CANDIDATES, tune.LAYOUT_CANDIDATES)
    tune.DEFAULT_CANDIDATES, tune.SPW_CANDIDATES = cands, [1, 2]
    tune.TAIL_CANDIDATES, tune.LAYOUT_CANDIDATES = [0, 1024], [0, 1]
    try:
        ps, shape, hit = tune.tuned_params(ctx, [q.qp for q in qps], base, M.FP64, path=path)
        assert not hit
        ps2, shape2, hit2 = tune.tuned_params(ctx, [q.qp for q in qps], base, M.FP64, path=path)
    finally:
        tune.DEFAULT_CANDIDATES, tune.SPW_CANDIDATES, tune.TAIL_CANDIDATES, tune.LAYOUT_CANDIDATES = saved
    assert hit2 and shape2 == shape
    with open(path) as f:
        cache = json.load(f)
    key = tune.batch_key([q.qp for q in qps], M.FP64)
    assert cache[key]["shape"] == shape
    tried = [s for s, _ in cache[key]["ms"]]
    # 3 shapes, 1 more slices-per-wave at the best, 1 more tail point, 1 more placement
    assert len(tried) == 6 and all(s in tried for s in cands)
    assert all(p.tiles == shape["tiles"] and p.graph_steps == shape["graph_steps"]
               and p.tail_tiles == shape.get("tail_tiles", 0) and p.xcd_layout == shape.get("xcd_layout", 0)
               and p.slices_per_wave == shape.get("slices_per_wave", 1) for p in ps)
    # each tried shape: converged, and within the two solves' error bounds of
    # the tuned shape's solution (|x - x*| <= error_bound for each)
    sols = {}
    for s in tried:
        for q in qps:
            q.qp.reset_x()
        rc, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=[tune.apply(M.Params.from_buffer_copy(p), s)
                                                                   for p in base])
        assert all(x["code"] == 0 and x["error_bound"] > 0 for x in st), s
        sols[json.dumps(s, sort_keys=True)] = ([q.qp.get_x() for q in qps], [x["error_bound"] for x in st])
    ref_x, ref_b = sols[json.dumps(shape, sort_keys=True)]
    for xs, bs in sols.values():
        for x, xr, b, br in zip(xs, ref_x, bs, ref_b):
            assert np.linalg.norm(x - xr) <= b + br


@pytest.mark.parametrize("spw", [2])
def test_slices_per_wave_batch_equals_alone(ctx, batch, spw):
    """With a fixed slices_per_wave (and tiles), a QP's result is the same bits
    alone and inside the batch, as for the other launch-shape parameters."""
    qps, ab = batch
    kw = dict(tiles=2, graph_steps=16, slices_per_wave=spw)
    for q in qps:
        q.qp.reset_x()
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per_qp(ab, M.FP64, **kw))
    assert all(s["code"] == 0 for s in st)
    xs = [q.qp.get_x() for q in qps]
    for i in (0, 7, 15):
        qps[i].qp.reset_x()
        rc, s1 = M.solve_batch(ctx, [qps[i].qp], per_qp=per_qp(ab[i:i + 1], M.FP64, **kw))
        assert rc == 0 and s1[0]["iterations"] == st[i]["iterations"]
        np.testing.assert_array_equal(qps[i].qp.get_x(), xs[i])


@pytest.mark.parametrize("spw", [1, 2])
def test_slices_per_wave_parity_with_the_oracle(ctx, spw):
    """FP64 bars against the oracle on config 1 (tol 1e-8): solution 1e-8
    relative, objective 1e-10."""
    import oracle
    from tests.common import scene_qp
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=20000, alpha_bar=ab)
    q = ctx.matrix_csr(A).qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    rc, st = q.solve(M.params(tol=1e-8, max_it=20000, alpha_bar=ab, tiles=1, slices_per_wave=spw))
    x = q.get_x()
    assert code == 0 and rc == 0
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
    f = lambda v: 0.5 * v @ (A @ v) - b @ v
    assert abs(f(x) - f(xr)) <= 1e-10 * abs(f(xr))
    with pytest.raises(M.MprgpError) as e:
        q.solve(M.params(tol=1e-8, max_it=10, alpha_bar=ab, slices_per_wave=3))
    assert e.value.code == -2 and "slices_per_wave" in M.last_error()
