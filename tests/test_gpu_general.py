"""GPU parity of the coupled-constraint solve (mprgp_general_solve:
MPRGPGeneralCon::solve with GeneralConProjector, MPRGPSolver.h:594-626,
MPRGPProjection.h:400-520) against the oracle's restatement.

Bars, as for the decoupled FP64 form: solution relative error <= 1e-8, the
same active rows except rows within 1e-10 of their bound, objective relative
error <= 1e-10; on tiny coupled QPs, the enumerated minimiser to 1e-8."""
import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import scene_qp
from tests.test_oracle_general import brute_qp, tiny_coupled_qp

pytestmark = pytest.mark.gpu


def qp_of(ctx, A, b, x0):
    q = ctx.matrix_csr(A).qp()
    q.set_rhs(b)
    q.set_x(x0)
    return q


def beam_with_contacts(seed=2, m=12):
    sc = problems.config1_flat()
    A, b, node, normal, c, x0 = scene_qp(sc)
    Jg = problems.rows_to_J(node, normal, sc.n_nodes)
    Jc, cc = problems.self_contact_rows(sc, np.random.default_rng(seed), m)
    return sc, A, b, sp.vstack([Jg, Jc]).tocsr(), np.concatenate([c, cc]), x0, len(c)


def test_coupled_rows_on_a_beam_match_the_oracle(ctx):
    sc, A, b, J, c, x0, m_ground = beam_with_contacts()
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr = oracle.general_solve(A, b, J, c, x0, tol=1e-9, max_it=3000, alpha_bar=ab)
    print(f"beam: {A.shape[0]} DoF, {J.shape[0]} rows; oracle {sr['iterations']} iterations", flush=True)
    q = qp_of(ctx, A, b, x0)
    rc, sg = M.general_solve(q, J, c, tol=1e-9, max_it=3000, alpha_bar=ab, project_first=1)
    xg = q.get_x()
    print(f"oracle {sr['iterations']} it ({sr['num_cg']}/{sr['num_exp']}/{sr['num_prop']}), "
          f"gpu {sg['iterations']} it ({sg['num_cg']}/{sg['num_exp']}/{sg['num_prop']}), "
          f"{sg['elapsed_ms']:.1f} ms")
    assert code == 0 and rc == 0
    assert sr["num_prop"] >= 1  # the path covers a proportioning step
    rel = np.linalg.norm(xg - xr) / np.linalg.norm(xr)
    assert rel <= 1e-8, rel
    f = lambda x: 0.5 * x @ (A @ x) - b @ x
    assert abs(f(xg) - f(xr)) <= 1e-10 * abs(f(xr))
    sg_, sr_ = J @ xg - c, J @ xr - c
    differ = (np.abs(sg_) < 1e-9) != (np.abs(sr_) < 1e-9)
    near = (np.abs(sg_) < 1e-10) | (np.abs(sr_) < 1e-10)
    assert not np.any(differ & ~near)
    assert np.any(np.abs(sr_[m_ground:]) < 1e-9), "no coupled row active"
    assert sg["residual"] <= 1e-9 and sg["error_bound"] == -1.0


@pytest.mark.parametrize("seed", [0, 1, 3, 6, 7])
def test_tiny_coupled_qps_reach_the_enumerated_minimiser(ctx, seed):
    A, b, J, c, x0 = tiny_coupled_qp(seed)
    q = qp_of(ctx, A, b, x0)
    rc, st = M.general_solve(q, J, c, tol=1e-11, max_it=3000, project_first=1)
    assert rc == 0, st
    np.testing.assert_allclose(q.get_x(), brute_qp(A, b, J, c), atol=1e-8)
    code, xr, sr = oracle.general_solve(A, b, J, c, x0, tol=1e-11, max_it=3000)
    assert abs(st["lambda_max"] - sr["lambda_max"]) <= 1e-10 * sr["lambda_max"]  # same power method


def test_decoupled_rows_reach_the_decoupled_minimiser(ctx):
    """With per-node rows only, the coupled solve (in-face Jacobi) and the
    decoupled solve (tangent Jacobi) reach the same minimiser."""
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    q = qp_of(ctx, A, b, x0)
    q.set_rows(node, normal, c)
    rc, sd = q.solve(tol=1e-10, max_it=20000, alpha_bar=ab, project_first=1)
    xd = q.get_x()
    q.reset_x()
    rc2, sgen = M.general_solve(q, J, c, tol=1e-10, max_it=2000, alpha_bar=ab, project_first=1)
    assert rc == rc2 == 0
    assert np.linalg.norm(q.get_x() - xd) <= 1e-8 * np.linalg.norm(xd)


def test_more_rows_than_one_dense_block(ctx):
    """J J^T is kept per connected component of its pattern, so the row count
    has no dense limit: 3,321 per-node rows (a 41x9x9-node beam) reach the
    decoupled solve's minimiser, and with 12 coupled rows on top the solve
    converges to a feasible point whose objective is not below the
    minimum without them (more rows, smaller feasible set).  At tol 1e-9 the
    coupled case stalls above its minimum, the reference's behaviour on
    overlapping rows (test_general_solve_can_stall_on_overlapping_rows)."""
    sc = problems.beam_scene("beam40x8x8", 40, 8, 8, seed=5, tol=1e-8)
    A, b, node, normal, c, x0 = scene_qp(sc)
    assert len(node) > 2048
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    q = qp_of(ctx, A, b, x0)
    q.set_rows(node, normal, c)
    rc, sd = q.solve(tol=1e-10, max_it=20000, project_first=1)
    xd = q.get_x()
    ab = sd["alpha_bar"]
    q.reset_x()
    rc2, sg = M.general_solve(q, J, c, tol=1e-10, max_it=20000, alpha_bar=ab, project_first=1)
    print(f"{len(node)} rows: decoupled {sd['iterations']} it, coupled path {sg['iterations']} it, "
          f"{sg['elapsed_ms']:.1f} ms", flush=True)
    assert rc == rc2 == 0
    assert np.linalg.norm(q.get_x() - xd) <= 1e-8 * np.linalg.norm(xd)
    Jc, cc = problems.self_contact_rows(sc, np.random.default_rng(2), 12)
    J2, c2 = sp.vstack([J, Jc]).tocsr(), np.concatenate([c, cc])
    q.reset_x()
    rc3, s3 = M.general_solve(q, J2, c2, tol=1e-6, max_it=20000, alpha_bar=ab, project_first=1)
    print(f"with 12 coupled rows: {s3['iterations']} it, code {rc3}", flush=True)
    x3 = q.get_x()
    assert rc3 == 0 and s3["residual"] <= 1e-6
    assert (J2 @ x3 - c2).min() >= -1e-9
    f = lambda x: 0.5 * x @ (A @ x) - b @ x
    assert f(x3) >= f(xd) - 1e-10 * abs(f(xd))


def test_too_many_rows_on_one_node_are_rejected(ctx):
    """More rows sharing a column than one component may hold (2048) are
    rejected before J J^T is formed, with the row named."""
    A, b, J, c, x0 = tiny_coupled_qp(0)
    rng = np.random.default_rng(9)
    m = 2100
    nrm = rng.normal(size=(m, 3))
    Jbig = sp.csr_matrix((nrm.ravel(), np.tile(np.arange(3), m), 3 * np.arange(m + 1)), shape=(m, A.shape[0]))
    q = qp_of(ctx, A, b, x0)
    with pytest.raises(M.MprgpError) as e:
        M.general_solve(q, Jbig, Jbig @ x0 - 1.0, tol=1e-8, max_it=10, alpha_bar=0.01)
    assert e.value.code == -2 and "shares a column with 2100 rows" in M.last_error()


def test_no_rows_is_unconstrained(ctx):
    """With no rows every projection, face and BETA is empty: the solve is
    preconditioned CG with the monotone test and reaches A^-1 b."""
    A, b, J, c, x0 = tiny_coupled_qp(0)
    J0 = sp.csr_matrix((0, A.shape[0]))
    code, xr, sr = oracle.general_solve(A, b, J0, np.zeros(0), x0, tol=1e-11, max_it=3000)
    q = qp_of(ctx, A, b, x0)
    rc, st = M.general_solve(q, J0, np.zeros(0), tol=1e-11, max_it=3000, project_first=1)
    assert code == 0 and rc == 0, st
    xs = np.linalg.solve(A.toarray(), b)
    assert np.abs(q.get_x() - xs).max() <= 1e-9 * np.abs(xs).max()
    assert np.abs(q.get_x() - xr).max() <= 1e-9 * np.abs(xs).max()
    assert st["active"] == 0 and st["residual"] <= 1e-11


def test_coupled_solve_errors(ctx):
    A, b, J, c, x0 = tiny_coupled_qp(0)
    q = qp_of(ctx, A, b, x0)
    with pytest.raises(M.MprgpError) as e:
        M.general_solve(q, J, c, precision=M.MIXED)
    assert e.value.code == -2
    Jd = J.toarray()
    # every row twice: the active rows' J_a J_a^T is singular (oracle code -2)
    twice = sp.csr_matrix(np.vstack([Jd, Jd]))
    code, _, _ = oracle.general_solve(A, b, twice, np.concatenate([c, c]), x0, tol=1e-8, max_it=100,
                                      alpha_bar=0.01, project_first=1)
    assert code == -2
    q.reset_x()
    with pytest.raises(M.MprgpError) as e:
        M.general_solve(q, twice, np.concatenate([c, c]), tol=1e-8, max_it=100, project_first=1, alpha_bar=0.01)
    assert e.value.code == -2 and "J J^T" in M.last_error()
    # a repeated row with a shifted offset: never both on their planes, so
    # nothing singular is factored; the oracle runs to max_it (code -1)
    dup = sp.csr_matrix(np.vstack([Jd, Jd[:1]]))
    cd = np.concatenate([c, c[:1] + 1.0])
    code, _, sr = oracle.general_solve(A, b, dup, cd, x0, tol=1e-8, max_it=100, alpha_bar=0.01, project_first=1)
    q.reset_x()
    rc, sg = M.general_solve(q, dup, cd, tol=1e-8, max_it=100, project_first=1, alpha_bar=0.01)
    assert code == -1 and sg["code"] == -1 and rc != 0 and sg["iterations"] == sr["iterations"] == 100
