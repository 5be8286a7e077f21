"""Pins of the oracle's ICA comparator (oracle/ica_oracle.c: ICASolver and its
ProjectedGaussSeidel, ICASolver.h:20-463), each against what the mathematics
fixes: the projected Gauss-Seidel reaches the brute-force minimiser of its
nonnegative QP; with no rows ICA is block Gauss-Seidel on A x = R, whose one
sweep is the block-lower-triangular solve (Da + La) x1 = R - Ua x0 and whose
limit is A^-1 R; with rows ICA reaches the enumerated minimiser of the QP and
its KKT multipliers."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle
from tests.common import random_spd
from tests.test_oracle_general import brute_nonneg, brute_qp, random_coupled


def block_dominant_spd(rng, n_nodes, couple=0.3):
    """Block SPD matrix with 3x3 diagonal blocks and sparse couplings, strictly
    diagonally dominant (block Gauss-Seidel converges)."""
    n = 3 * n_nodes
    A = np.zeros((n, n))
    for v in range(n_nodes):
        A[3 * v:3 * v + 3, 3 * v:3 * v + 3] = random_spd(rng, 3, cond=3.0) + 3.0 * np.eye(3)
    for _ in range(2 * n_nodes):
        i, j = rng.choice(n_nodes, 2, replace=False)
        Bk = couple * rng.normal(size=(3, 3))
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += Bk
        A[3 * j:3 * j + 3, 3 * i:3 * i + 3] += Bk.T
    d = np.abs(A).sum(1) - np.abs(np.diag(A))
    A += np.diag(np.maximum(0.0, d - np.diag(A)) + 0.5)
    return A


def test_inverse_blocks():
    rng = np.random.default_rng(1)
    for _ in range(20):
        M = random_spd(rng, 3, cond=100.0)
        np.testing.assert_allclose(oracle.inv3(M) @ M, np.eye(3), atol=1e-12)


@pytest.mark.parametrize("seed", range(8))
def test_pgs_reaches_the_nonnegative_minimiser(seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, 8))
    B = random_spd(rng, m, cond=20.0)
    B += np.diag(np.abs(B).sum(1))  # diagonally dominant: PGS converges
    c = rng.normal(size=m)
    met, lam, it, res = oracle.pgs(sp.csr_matrix(B), c, max_it=100000, tol=1e-15)
    assert met and res <= 1e-15
    np.testing.assert_allclose(lam, brute_nonneg(B, c), atol=1e-12)


def test_pgs_counts_like_the_reference():
    """`iterations` is the sweep index at the stop: a lambda already at the
    solution stops after sweep 0; max_it when the tolerance is never met."""
    B = sp.csr_matrix(np.array([[2.0, 0.5], [0.5, 3.0]]))
    c = np.array([1.0, -1.0])
    met, lam, it, _ = oracle.pgs(B, c, max_it=1000, tol=1e-14)
    met2, lam2, it2, _ = oracle.pgs(B, c, lam0=lam, max_it=1000, tol=1e-14)
    assert met and met2 and it2 == 0 and np.array_equal(lam, lam2)
    met3, _, it3, _ = oracle.pgs(B, c, max_it=1, tol=0.0)
    assert not met3 and it3 == 1


def test_one_sweep_is_the_block_lower_solve():
    rng = np.random.default_rng(2)
    A = block_dominant_spd(rng, 6)
    R = rng.normal(size=18)
    x0 = rng.normal(size=18)
    code, x1, _, st = oracle.ica_solve(sp.csr_matrix(A), R, None, np.zeros(0), x0, tol=0.0, max_it=1)
    assert code == -1 and st["iterations"] == 1
    L = np.zeros_like(A)
    U = np.zeros_like(A)
    for i in range(18):
        for j in range(18):
            if i // 3 >= j // 3:
                L[i, j] = A[i, j]  # Da + La
            else:
                U[i, j] = A[i, j]  # Ua
    # L is block lower triangular (its diagonal blocks are full): a dense solve
    np.testing.assert_allclose(x1, np.linalg.solve(L, R - U @ x0), rtol=1e-12, atol=1e-12)


def test_no_rows_is_block_gauss_seidel_to_the_solution():
    rng = np.random.default_rng(3)
    A = block_dominant_spd(rng, 10)
    R = rng.normal(size=30)
    code, x, _, st = oracle.ica_solve(sp.csr_matrix(A), R, None, np.zeros(0), np.zeros(30), tol=1e-12,
                                      max_it=10000)
    assert code == 0 and st["residual"] <= 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, R), atol=1e-11)


def kkt_multipliers(A, b, J, c, x):
    Jd = J.toarray()
    act = np.abs(Jd @ x - c) < 1e-8
    lam = np.zeros(len(c))
    if act.any():
        lam[act] = np.linalg.lstsq(Jd[act].T, A @ x - b, rcond=None)[0]
    return lam


@pytest.mark.parametrize("seed", range(6))
def test_ica_reaches_the_minimiser_with_coupled_rows(seed):
    rng = np.random.default_rng(100 + seed)
    n_nodes = 6
    A = block_dominant_spd(rng, n_nodes)
    J = random_coupled(rng, n_nodes=n_nodes, m=4)
    x0 = rng.normal(size=3 * n_nodes)
    b = A @ x0 + rng.normal(size=3 * n_nodes)  # pulls some rows onto their planes
    c = J @ x0 - 0.1 * rng.random(4)
    code, x, lam, st = oracle.ica_solve(sp.csr_matrix(A), b, J, c, x0, tol=1e-11, max_it=20000,
                                        gs_max_it=200, gs_tol=1e-14)
    assert code == 0, st
    xb = brute_qp(A, b, J, c)
    np.testing.assert_allclose(x, xb, atol=1e-8)
    np.testing.assert_allclose(lam, kkt_multipliers(A, b, J, c, xb), atol=1e-7)
    assert np.all(lam >= 0.0)


@pytest.mark.parametrize("seed", range(4))
def test_ica_reaches_the_minimiser_with_decoupled_rows(seed):
    """One row per constrained node, as the plane collider makes them: B =
    J Da^-1 J^T is diagonal and one PGS sweep solves it."""
    rng = np.random.default_rng(200 + seed)
    n_nodes = 8
    A = block_dominant_spd(rng, n_nodes)
    nodes = rng.choice(n_nodes, 4, replace=False)
    rows = np.zeros((4, 3 * n_nodes))
    for i, v in enumerate(nodes):
        nv = rng.normal(size=3)
        rows[i, 3 * v:3 * v + 3] = nv / np.linalg.norm(nv)
    J = sp.csr_matrix(rows)
    x0 = rng.normal(size=3 * n_nodes)
    b = A @ x0 - 2.0 * J.T @ np.ones(4)
    c = J @ x0
    code, x, lam, st = oracle.ica_solve(sp.csr_matrix(A), b, J, c, x0, tol=1e-11, max_it=20000)
    assert code == 0, st
    np.testing.assert_allclose(x, brute_qp(A, b, J, c), atol=1e-8)


def test_b_is_j_dinv_jt():
    rng = np.random.default_rng(5)
    n_nodes = 7
    A = block_dominant_spd(rng, n_nodes)
    J = random_coupled(rng, n_nodes=n_nodes, m=5)
    dinv = np.stack([np.linalg.inv(A[3 * v:3 * v + 3, 3 * v:3 * v + 3]) for v in range(n_nodes)])
    B = oracle.ica_b(J, n_nodes, dinv)
    np.testing.assert_allclose(B.toarray(), J.toarray() @ sla.block_diag(*dinv) @ J.toarray().T, atol=1e-13)
    assert np.all(np.diff(B.indptr) > 0) and B.has_sorted_indices
