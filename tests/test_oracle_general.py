"""Pins of the oracle's coupled-constraint solve (general_oracle.c):
GeneralConProjector (MPRGPProjection.h:400-520) and MPRGPGeneralCon::solve
(MPRGPSolver.h:594-626), checked against brute-force enumeration of active
sets, KKT conditions and the decoupled special case."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from mprgp_amd import problems
from tests.common import coupled_rows, random_spd, scene_qp


def brute_nonneg(G, h):
    m = len(h)
    for k in range(m + 1):
        for S in itertools.combinations(range(m), k):
            lam = np.zeros(m)
            if S:
                lam[list(S)] = np.linalg.solve(G[np.ix_(S, S)], h[list(S)])
            w = h - G @ lam
            if np.all(lam >= -1e-12) and np.all(w <= 1e-10 * max(1.0, np.abs(h).max())):
                return lam
    raise AssertionError("no KKT point")


def brute_project(J, c, x):
    """Euclidean projection onto {J y >= c} by enumeration of active rows."""
    J = J.toarray() if sp.issparse(J) else J
    best = None
    for k in range(J.shape[0] + 1):
        for S in itertools.combinations(range(J.shape[0]), k):
            S = list(S)
            y = x.copy()
            mu = np.zeros(0)
            if S:
                JS = J[S]
                mu = np.linalg.solve(JS @ JS.T, c[S] - JS @ x)
                y = x + JS.T @ mu
            if np.all(mu >= -1e-12) and np.all(J @ y >= c - 1e-10):
                d = np.linalg.norm(y - x)
                if best is None or d < best[0] - 1e-14:
                    best = (d, y)
    return best[1]


def brute_qp(A, b, J, c):
    """Exact minimiser of 1/2 x'Ax - b'x s.t. J x >= c (enumeration of active rows)."""
    A = A.toarray() if sp.issparse(A) else A
    J = J.toarray() if sp.issparse(J) else J
    n, m = A.shape[0], J.shape[0]
    best = None
    for k in range(m + 1):
        for S in itertools.combinations(range(m), k):
            S = list(S)
            K = np.zeros((n + k, n + k))
            K[:n, :n] = A
            K[:n, n:] = -J[S].T
            K[n:, :n] = J[S]
            sol = np.linalg.solve(K, np.concatenate([b, c[S]]))
            x, mu = sol[:n], sol[n:]
            if np.all(J @ x >= c - 1e-10) and np.all(mu >= -1e-10):
                f = 0.5 * x @ A @ x - b @ x
                if best is None or f < best[1] - 1e-14:
                    best = (x, f)
    return best[0]


@pytest.mark.parametrize("seed", range(8))
def test_qp_nonneg_is_the_enumerated_minimiser(seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, 8))
    G = random_spd(rng, m, cond=1e3)
    h = rng.normal(size=m)
    rc, lam = oracle.qp_nonneg(G, h)
    assert rc == 0
    np.testing.assert_allclose(lam, brute_nonneg(G, h), atol=1e-10 * max(1, np.abs(h).max()))


def random_coupled(rng, n_nodes=6, m=4):
    """m rows of 12 nonzeros (a vertex against a triangle, LinearConCollider.cpp:
    186-217) over n_nodes nodes, full row rank."""
    while True:
        rows = []
        for _ in range(m):
            i, j, k, l = rng.choice(n_nodes, 4, replace=False)
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            w = rng.dirichlet([2, 2, 2])
            r = np.zeros(3 * n_nodes)
            r[3 * i:3 * i + 3] += nrm
            for node, wt in zip((j, k, l), w):
                r[3 * node:3 * node + 3] -= wt * nrm
            rows.append(r)
        J = np.array(rows)
        if np.linalg.matrix_rank(J) == m:
            return sp.csr_matrix(J)


@pytest.mark.parametrize("seed", range(6))
def test_projection_is_the_nearest_feasible_point(seed):
    rng = np.random.default_rng(seed)
    J = random_coupled(rng)
    x = rng.normal(size=J.shape[1])
    c = J @ x + rng.normal(size=J.shape[0]) * 0.5  # some rows violated
    P = oracle.GeneralProjector(J, c)
    y = P.project(x)
    np.testing.assert_allclose(y, brute_project(J, c, x), atol=1e-10)
    assert P.is_feasible(y)
    np.testing.assert_allclose(P.project(y), y, atol=1e-12)  # idempotent


@pytest.mark.parametrize("seed", range(6))
def test_phi_and_beta_on_the_active_rows(seed):
    rng = np.random.default_rng(seed)
    J = random_coupled(rng, n_nodes=8, m=5)
    x = rng.normal(size=J.shape[1])
    Jd = J.toarray()
    act = rng.random(J.shape[0]) < 0.6
    c = Jd @ x + np.where(act, 0.0, -1.0)  # active rows exactly on their plane
    P = oracle.GeneralProjector(J, c)
    assert P.decide_face(x) == 0
    assert np.array_equal(P.face().astype(bool), act)
    g = rng.normal(size=J.shape[1])
    phi = P.phi(g)
    Ja = Jd[act]
    if Ja.shape[0]:
        np.testing.assert_allclose(Ja @ phi, 0.0, atol=1e-12)
        ref = g - Ja.T @ np.linalg.solve(Ja @ Ja.T, Ja @ g)
        np.testing.assert_allclose(phi, ref, atol=1e-12)
    else:
        np.testing.assert_array_equal(phi, g)
    beta = P.beta(g, phi)
    gp = g - phi
    if Ja.shape[0]:
        lam = np.linalg.lstsq(Ja.T, gp - beta, rcond=None)[0]
        np.testing.assert_allclose(Ja.T @ lam, gp - beta, atol=1e-12)
        assert np.all(lam >= -1e-12)                      # multipliers
        assert np.all(Ja @ beta <= 1e-12)                 # KKT of BETA's QP
        np.testing.assert_allclose(lam * (Ja @ beta), 0.0, atol=1e-12)
    ab = 0.3
    px = P.project(x - ab * phi)
    assert P.phitphi(x, ab, phi) == pytest.approx(max(0.0, (x - px) @ phi / ab), rel=1e-12, abs=1e-300)


def test_decoupled_rows_reduce_to_the_decoupled_projector():
    rng = np.random.default_rng(7)
    n_nodes = 10
    node = np.array([0, 3, 4, 8])
    nrm = rng.normal(size=(4, 3))
    J = problems.rows_to_J(node, nrm, n_nodes)
    x = rng.normal(size=3 * n_nodes)
    c = J @ x + rng.normal(size=4)
    G = oracle.GeneralProjector(J, c)
    D = oracle.Projector(J, c)
    np.testing.assert_allclose(G.project(x), D.project(x), atol=1e-13)
    y = G.project(x)
    G.decide_face(y)
    D.decide_face(y)
    assert np.array_equal(G.face().astype(bool), D.face().astype(bool))
    g = rng.normal(size=3 * n_nodes)
    np.testing.assert_allclose(G.phi(g), D.phi(g), atol=1e-13)
    np.testing.assert_allclose(G.beta(g, G.phi(g)), D.beta(g, D.phi(g)), atol=1e-13)


def tiny_coupled_qp(seed):
    rng = np.random.default_rng(seed)
    n_nodes = 5
    A = sp.csr_matrix(random_spd(rng, 3 * n_nodes, cond=50.0))
    J = random_coupled(rng, n_nodes=n_nodes, m=4)  # 4 rows of 4 nodes each over 5 nodes
    b = rng.normal(size=3 * n_nodes)
    x0 = rng.normal(size=3 * n_nodes)
    c = J @ x0 - 0.2  # x0 feasible; the minimiser presses on some rows
    return A, b, J, c, x0


@pytest.mark.parametrize("seed", [0, 1, 3, 6, 7])
def test_general_solve_finds_the_enumerated_minimiser(seed):
    A, b, J, c, x0 = tiny_coupled_qp(seed)
    code, x, st = oracle.general_solve(A, b, J, c, x0, tol=1e-11, max_it=20000)
    assert code == 0, st
    np.testing.assert_allclose(x, brute_qp(A, b, J, c), atol=1e-8)


def test_general_solve_can_stall_on_overlapping_rows():
    """Reference behaviour kept: with rows sharing nodes (J J^T not
    diagonal) the expansion step x = P(x - abar phi) (MPRGPSolver.h:250-271)
    projects onto coupled rows and can raise f; then CG steps are rejected by
    the monotone test and expansion repeats.  On this instance the solve
    stalls at f above the minimum until max_it (code -1)."""
    A, b, J, c, x0 = tiny_coupled_qp(2)
    code, x, st = oracle.general_solve(A, b, J, c, x0, tol=1e-9, max_it=2000)
    assert code == -1 and st["num_exp"] > 1900
    f = lambda v: 0.5 * v @ (A @ v) - b @ v
    xb = brute_qp(A, b, J, c)
    assert f(x) > f(xb) + 0.1
    assert oracle.GeneralProjector(J, c).is_feasible(x)


def test_general_solve_on_a_beam_with_coupled_rows():
    """Ground rows plus vertex-triangle rows inside the beam: feasible,
    nonnegative multipliers, complementarity, tangential gradient zero."""
    sc = problems.config1_flat()
    A, b, node, normal, c, x0 = scene_qp(sc)
    Jg = problems.rows_to_J(node, normal, sc.n_nodes)
    Jc, cc = coupled_rows(sc, np.random.default_rng(2), m=12)
    J = sp.vstack([Jg, Jc]).tocsr()
    cv = np.concatenate([c, cc])
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, x, st = oracle.general_solve(A, b, J, cv, x0, tol=1e-9, max_it=20000, alpha_bar=ab)
    assert code == 0, st
    Jd = J.toarray()
    s = Jd @ x - cv
    assert s.min() >= -1e-9
    g = A @ x - b
    act = np.abs(s) < 1e-8
    mu = np.linalg.lstsq(Jd[act].T, g, rcond=None)[0]
    assert mu.min() >= -1e-7
    assert np.linalg.norm(g - Jd[act].T @ mu) <= 1e-7 * max(1.0, np.linalg.norm(g))
    # the coupled rows matter: some are active at the solution
    assert np.any(act[len(c):]), "no coupled row active"
    # decoupled rows only: the general solve reaches the decoupled solve's minimiser
    code_d, xd, _, _ = oracle.decoupled_solve(A, b, Jg, c, x0, tol=1e-10, max_it=20000, alpha_bar=ab)
    code_g, xg, _ = oracle.general_solve(A, b, Jg, c, x0, tol=1e-10, max_it=20000, alpha_bar=ab)
    assert code_d == code_g == 0
    assert np.linalg.norm(xg - xd) <= 1e-8 * np.linalg.norm(xd)
