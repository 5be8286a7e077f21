"""What the library reports besides x: per-op launch counts (the byte model's
input), the eigenvector the power method keeps (specRad's ev,
MPRGPSolver.h:42-79), and the context's device-memory account."""
import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
from mprgp_amd import problems, workloads
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


def test_op_counts_cover_every_launch(ctx):
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    q = ctx.matrix_csr(A).qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    rc, st = q.solve(M.params(tol=1e-8, max_it=3000, project_first=1))
    assert rc == 0
    counts = q.op_counts()
    op = workloads.OPS.index
    assert int(counts.sum()) == st["steps"]  # every launch the QP ran counts once, under its op
    assert int(counts[op("POW_IT")]) >= st["power_iters"] >= 1
    assert int(counts[op("EXP3")]) == st["num_exp"] >= 1  # an expansion ends in EXP3
    assert int(counts[op("PROP2")]) == st["num_prop"]
    assert int(counts[op("FINAL")]) == 1


def test_ev_after_an_early_stop(ctx):
    """The power method keeps ev when its difference test stops it early:
    unit length, and A ev = lambda ev for the lambda it reports."""
    rng = np.random.default_rng(0)
    n = 15
    d = np.arange(1.0, n + 1.0)
    d[-1] = 30.0  # the top eigenvalue well apart from the next (14)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = (Q * d) @ Q.T
    A = 0.5 * (A + A.T)
    q = ctx.matrix_csr(sp.csr_matrix(A)).qp()
    q.set_rhs(rng.normal(size=n))
    q.set_x(np.zeros(n))
    assert q.get_ev() is None
    rc, st = q.solve(M.params(tol=1e-8, max_it=0, power_tol=1e-10))
    assert rc == 0 and st["ev_valid"] == 1 and st["power_iters"] < 1000
    ev = q.get_ev()
    lam = st["lambda_max"]
    assert ev is not None and abs(np.linalg.norm(ev) - 1.0) <= 1e-9
    assert abs(lam - 30.0) <= 1e-8 * 30.0
    assert np.linalg.norm(A @ ev - lam * ev) <= 1e-8 * lam


def test_memory_account_follows_allocations(ctx):
    used0, budget = ctx.memory()
    sc = problems.config1()
    A, *_ = scene_qp(sc)
    m = ctx.matrix_csr(A)
    used1, _ = ctx.memory()
    assert used1 - used0 >= m.info()["bytes"] > 0
    m.close()
    assert ctx.memory()[0] == used0
