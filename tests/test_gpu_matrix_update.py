"""New values on a matrix's pattern (mprgp_matrix_update_bsr3: the values of
A change between Newton iterations and frames, the pattern does not, as in
MprgpFemSolver.cpp's buildLinearSystem each step) must give the matrix a
fresh upload of the same blocks gives, bit for bit: the SELL values, the
padding slots, D^-1 and every solve after it."""
import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
from mprgp_amd import problems
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


def block_csr(A):
    B = sp.bsr_matrix(A, blocksize=(3, 3))
    B.sort_indices()
    return B.indptr.astype(np.int32), B.indices.astype(np.int32), np.ascontiguousarray(B.data)


def scaled_blocks(indptr, indices, data, d):
    """D A D on the same pattern, block by block (SPD stays SPD)."""
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    dr = d.reshape(-1, 3)[rows]
    dc = d.reshape(-1, 3)[indices]
    return data * dr[:, :, None] * dc[:, None, :]


def solve_on(m, b, node, normal, c, x0):
    q = m.qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    rc, st = q.solve(M.params(tol=1e-8, max_it=3000, project_first=1))
    return rc, st, q.get_x()


def test_update_equals_a_fresh_upload(ctx):
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    indptr, indices, data = block_csr(A)
    n_nodes = sc.n_nodes
    rng = np.random.default_rng(3)
    new = scaled_blocks(indptr, indices, data, 1.0 + 0.5 * rng.random(3 * n_nodes))
    m1 = ctx.matrix_bsr3(n_nodes, indptr, indices, data)
    _, _, x_old = solve_on(m1, b, node, normal, c, x0)  # a solve on the old values first
    m1.update_bsr3(new)
    m2 = ctx.matrix_bsr3(n_nodes, indptr, indices, new)
    x = rng.normal(size=3 * n_nodes)
    assert np.array_equal(m1.spmv(x), m2.spmv(x))
    assert np.array_equal(m1.diagonal(), m2.diagonal())
    A2 = sp.bsr_matrix((new, indices, indptr), shape=A.shape).tocsr()
    np.testing.assert_allclose(m1.spmv(x), A2 @ x, rtol=0, atol=1e-12 * np.abs(A2 @ x).max())
    rc1, st1, x1 = solve_on(m1, b, node, normal, c, x0)
    rc2, st2, x2 = solve_on(m2, b, node, normal, c, x0)
    assert rc1 == rc2 == 0
    assert np.array_equal(x1, x2) and st1["iterations"] == st2["iterations"]
    assert not np.array_equal(x1, x_old)


def test_non_finite_update_is_rejected_and_keeps_the_values(ctx):
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    indptr, indices, data = block_csr(A)
    m = ctx.matrix_bsr3(sc.n_nodes, indptr, indices, data)
    x = np.random.default_rng(4).normal(size=3 * sc.n_nodes)
    y0 = m.spmv(x)
    for where in (7, len(data) // 2, len(data) - 1):
        bad = data * 2.0
        bad[where, 1, 2] = np.nan
        bad[len(data) - 1, 0, 0] = np.inf if where != len(data) - 1 else bad[len(data) - 1, 0, 0]
        with pytest.raises(M.MprgpError) as e:
            m.update_bsr3(bad)
        assert e.value.code == -2
        assert M.last_error().endswith(f"at {9 * where + 5}")  # the first non-finite value, in block order
        assert np.array_equal(m.spmv(x), y0)
