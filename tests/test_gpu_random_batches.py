"""Seeded random batches through the batch driver: for random launch shapes
(wavefronts per workgroup, launches per graph replay, tail switch point, XCD
placement, slices per wavefront), random mixes of FP64 and mixed-form QPs and
QPs that finish at different launches (so the driver drops finished QPs'
tiles while replays are in flight), every QP's result is the same, bit for
bit, alone and in its batch, and with 16-bit and 32-bit column indices."""
import numpy as np
import pytest

import mprgp_amd as M
from mprgp_amd import problems, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", range(12))
def test_random_batch_equals_alone(ctx, seed, monkeypatch):
    rng = np.random.default_rng(1000 + seed)
    k = int(rng.integers(2, 20))
    scenes = problems.batch_scenes(k, seed=int(rng.integers(1 << 30)), dof_range=(1.5e3, 2.5e4), tol=1e-6,
                                   max_it=4000)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=1e-6, max_it=0, power_tol=1e-6))
    ab = [s["alpha_bar"] for s in st]
    prec = [M.MIXED if rng.random() < 0.6 else M.FP64 for _ in qps]
    kw = dict(tiles=int(rng.choice([1, 2, 4])), graph_steps=int(rng.choice([8, 16, 32])),
              tail_tiles=int(rng.choice([0, 64, 1 << 30])), xcd_layout=int(rng.integers(0, 2)),
              slices_per_wave=int(rng.choice([1, 1, 2])))
    per = [M.params(tol=1e-6, max_it=4000, alpha_bar=a, project_first=1, precision=p, from_x0=1, **kw)
           for a, p in zip(ab, prec)]
    print(f"{k} QPs, {kw}, mixed {sum(p == M.MIXED for p in prec)}", flush=True)
    rc, sb = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
    assert rc == 0 and all(s["code"] == 0 for s in sb)
    print("iterations", sorted(s["iterations"] for s in sb), flush=True)
    xb = [q.qp.get_x() for q in qps]
    monkeypatch.setenv("MPRGP_INDEX32", "1")
    rc, s32 = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
    for q, x, a, b in zip(qps, xb, sb, s32):
        np.testing.assert_array_equal(q.qp.get_x(), x)
        assert a["iterations"] == b["iterations"]
    monkeypatch.setenv("MPRGP_INDEX32", "0")
    for i in rng.choice(k, size=min(3, k), replace=False):
        rc, s1 = M.solve_batch(ctx, [qps[i].qp], per_qp=[per[i]])
        np.testing.assert_array_equal(qps[i].qp.get_x(), xb[i])
        assert s1[0]["iterations"] == sb[i]["iterations"]
