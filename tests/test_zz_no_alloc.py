"""mprgp_qp_solve_noalloc through the C ABI against mprgp_qp_solve.

The no-alloc entry point runs the same kernel with the same launch shape and
op sequence as mprgp_qp_solve, with its launch tables in the library's static
per-GPU scratch array; its results must equal mprgp_qp_solve's bit for bit
(solution, iteration and step counts, active count, residual, alpha_bar).
Two calls in a row on the same GPU, the second on a smaller QP, show the
shared scratch keeps nothing from the first call that matters; both forms
(FP64 and mixed) and the power method's start vector are covered, and a QP
beyond the tables' limit is rejected with MPRGP_EBADINPUT before any work."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
from mprgp_amd import _lib, problems
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


def make(ctx, scene):
    A, b, node, normal, c, x0 = scene_qp(scene)
    q = ctx.matrix_csr(A).qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    return q


def noalloc(q, p):
    st = M.Stats()
    rc = _lib.load().mprgp_qp_solve_noalloc(q.h, C.byref(p), C.byref(st))
    return rc, st.as_dict()


KEYS = ("code", "iterations", "num_cg", "num_exp", "num_prop", "active", "power_iters", "refinements",
        "fp64_finish", "steps", "residual", "alpha_bar", "lambda_max")


def solve_both(q, **kw):
    p = M.params(**kw)
    q.reset_x()
    rc_a, st_a = q.solve(p)
    x_a = q.get_x()
    q.reset_x()
    rc_b, st_b = noalloc(q, p)
    x_b = q.get_x()
    assert rc_a == rc_b, (rc_a, rc_b, M.last_error())
    for k in KEYS:
        assert st_a[k] == st_b[k], (k, st_a[k], st_b[k])
    assert np.array_equal(x_a, x_b)
    return st_b


def test_noalloc_equals_solve_on_two_calls_in_a_row(ctx):
    big = make(ctx, problems.batch_scenes(1, seed=21, dof_range=(3e4, 3e4))[0])
    small = make(ctx, problems.config1())
    assert big.n > small.n
    for q in (big, small):
        # FP64 from the power method's own start vector, then with alpha_bar given
        st = solve_both(q, tol=1e-8, max_it=5000, project_first=1, use_ev=0)
        assert st["code"] == 0 and st["power_iters"] > 0
        ab = st["alpha_bar"]
        solve_both(q, tol=1e-8, max_it=5000, project_first=1, alpha_bar=ab)
        # mixed: its FP32 storage exists after the first solve_both's mprgp_qp_solve
        st = solve_both(q, tol=1e-6, max_it=5000, project_first=1, alpha_bar=ab, precision=M.MIXED)
        assert st["code"] == 0 and st["refinements"] >= 1
        # two slices per wavefront: tiles of twice the rows, the same on both paths
        solve_both(q, tol=1e-8, max_it=5000, project_first=1, alpha_bar=ab, tiles=1, slices_per_wave=2)


def test_noalloc_limits_are_reported(ctx):
    # 2^20 + 64 nodes at one wavefront per workgroup: 16385 tiles, past the tables
    n = (1 << 20) + 64
    rowptr = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    blocks = np.tile(np.eye(3).ravel(), n)
    q = ctx.matrix_bsr3(n, rowptr, col, blocks).qp()
    q.set_rhs(np.ones(3 * n))
    q.set_x(np.zeros(3 * n))
    rc, _ = noalloc(q, M.params(tol=1e-6, max_it=10, alpha_bar=1.0, tiles=1))
    assert rc == _lib.EBADINPUT and "no-alloc tables hold 16384" in M.last_error(), M.last_error()
    # a mixed solve before the QP has FP32 storage
    small = make(ctx, problems.config1())
    rc, _ = noalloc(small, M.params(tol=1e-6, alpha_bar=1.0, precision=M.MIXED))
    assert rc == _lib.ESTATE and "FP32 storage" in M.last_error(), M.last_error()
