"""Pins the CPU oracle's solver functions (CPU only).

Each check is independent of the oracle's formulas: closed forms, the
characterising optimality conditions, brute force on tiny inputs, dense
linear algebra, and the reference's own test cases (test_case/MPRGPTest.cpp)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from mprgp_amd import problems
from tests.common import brute_force_qp, kkt_report, random_spd, scene_qp

EPS = 1e-9


def rand_rows(rng, N, m):
    node = rng.choice(N, size=m, replace=False).astype(np.int32)
    normal = rng.normal(size=(m, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    c = rng.normal(size=m)
    return node, normal, c


def test_spmv_and_func_value_match_dense():
    rng = np.random.default_rng(0)
    M = sp.random(60, 60, density=0.2, random_state=1, format="csr")
    A = (M + M.T + 10 * sp.eye(60)).tocsr()
    x, b = rng.normal(size=60), rng.normal(size=60)
    np.testing.assert_allclose(oracle.spmv(A, x), A.toarray() @ x, rtol=1e-14, atol=1e-13)
    f = 0.5 * x @ (A.toarray() @ x) - x @ b
    assert oracle.func_value(A, x, b) == pytest.approx(f, rel=1e-14)


def test_project_is_the_euclidean_projection():
    rng = np.random.default_rng(1)
    N, m = 40, 25
    node, normal, c = rand_rows(rng, N, m)
    J = problems.rows_to_J(node, normal, N)
    P = oracle.Projector(J, c)
    assert P.is_decoupled()
    for _ in range(5):
        x = rng.normal(size=3 * N) * 2
        y = P.project(x)
        X, Y = x.reshape(-1, 3), y.reshape(-1, 3)
        s = np.einsum("ij,ij->i", normal, Y[node]) - c
        assert s.min() >= -1e-12                                 # feasible
        t = np.einsum("ij,ij->i", normal, Y[node] - X[node])     # y - x = t n, t >= 0
        np.testing.assert_allclose(Y[node] - X[node], normal * t[:, None], atol=1e-12)
        assert t.min() >= -1e-12
        assert np.abs(t * s).max() <= 1e-12                      # complementarity
        free = np.setdiff1d(np.arange(N), node)
        np.testing.assert_array_equal(Y[free], X[free])
        np.testing.assert_allclose(P.project(y), y, atol=1e-13)  # idempotent


def test_decide_face_and_feasibility_thresholds():
    # one node, normal e_y, c = 0: n.x at -2e-9, -5e-10, 0, 5e-10, 2e-9
    J = problems.rows_to_J(np.array([0]), np.array([[0.0, 1.0, 0.0]]), 1)
    P = oracle.Projector(J, np.array([0.0]))
    for yv, face, feas in [(-2e-9, 0, False), (-5e-10, 1, True), (0.0, 1, True),
                           (5e-10, 1, True), (2e-9, 0, True)]:
        x = np.array([0.3, yv, -0.1])
        assert P.decide_face(x)[0] == face
        assert P.num_active() == face
        assert P.is_feasible(x) == feas


def test_phi_beta_split_invariants():
    rng = np.random.default_rng(2)
    N, m = 30, 20
    node, normal, c = rand_rows(rng, N, m)
    P = oracle.Projector(problems.rows_to_J(node, normal, N), c)
    face = (rng.uniform(size=m) < 0.5).astype(np.int8)
    P.set_face(face)
    g = rng.normal(size=3 * N)
    phi = P.phi(g)
    beta = P.beta(g, phi)
    G, PH, BE = g.reshape(-1, 3), phi.reshape(-1, 3), beta.reshape(-1, 3)
    act = node[face == 1]
    nact = normal[face == 1]
    ng = np.einsum("ij,ij->i", nact, G[act])
    # phi: tangential part of g on active nodes
    np.testing.assert_allclose(np.einsum("ij,ij->i", nact, PH[act]), 0, atol=1e-13)
    np.testing.assert_allclose(PH[act], G[act] - nact * ng[:, None], atol=1e-13)
    # beta: n min(0, n.g) on active nodes, zero elsewhere
    np.testing.assert_allclose(BE[act], nact * np.minimum(ng, 0)[:, None], atol=1e-13)
    others = np.setdiff1d(np.arange(N), act)
    np.testing.assert_array_equal(PH[others], G[others])
    np.testing.assert_allclose(BE[others], 0, atol=0)


def test_beta_with_foreign_phi_keeps_reference_quirk():
    # MPRGPProjection.h:368-383: rows outside the face subtract n (n.beta)
    # unscaled -- visible when BETA is handed a phi that is not PHI(g), as the
    # reference's own test_projector does.  Normal (0, 2, 0), inactive row.
    J = problems.rows_to_J(np.array([0]), np.array([[0.0, 2.0, 0.0]]), 2)
    P = oracle.Projector(J, np.array([0.0]))
    P.set_face(np.array([1], np.int8))   # make num_active > 0 ...
    P.set_face(np.array([0], np.int8))   # ... then the row inactive: num_active == 0 path
    g = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    phi = np.array([0.5, 0.0, 0.0, 2.0, 2.0, 2.0])
    np.testing.assert_allclose(P.beta(g, phi), g - phi)
    # two rows, one active: the inactive one takes the unscaled path
    J2 = problems.rows_to_J(np.array([0, 1]), np.array([[0.0, 2.0, 0.0], [0.0, 2.0, 0.0]]), 2)
    P2 = oracle.Projector(J2, np.array([0.0, 0.0]))
    P2.set_face(np.array([1, 0], np.int8))
    b = P2.beta(g, phi)
    # node 1: beta = g - phi = 0 -> unchanged; node 0: beta0 = (0.5, 1, 1),
    # lambda = max(0, 2*1/4) = 0.5, beta0 -= (0, 1, 0)
    np.testing.assert_allclose(b, [0.5, 0.0, 1.0, 0.0, 0.0, 0.0])
    g2 = np.array([0.0, 0.0, 0.0, 0.0, 3.0, 0.0])
    b2 = P2.beta(g2, np.zeros(6))
    # node 1 inactive: lambda = n.beta = 6 (unscaled), beta1 -= n*6 -> -9
    np.testing.assert_allclose(b2, [0, 0, 0, 0, 3.0 - 12.0, 0])


def test_phitphi_closed_forms():
    J = problems.rows_to_J(np.array([0]), np.array([[0.0, 1.0, 0.0]]), 1)
    P = oracle.Projector(J, np.array([0.0]))
    x = np.array([0.0, 1.0, 0.0])
    phi = np.array([0.3, 0.4, 0.0])
    ab = 2.0
    # unconstrained move (stays feasible): ((ab phi).phi)/ab = |phi|^2
    assert P.phitphi(x, ab, phi) == pytest.approx(0.25, rel=1e-15)
    # blocked: x - ab*phi has y = 1 - 2*1 = -1 -> projected to 0; x - px = (0.6, 1, 0)
    phi2 = np.array([0.3, 1.0, 0.0])
    assert P.phitphi(x, ab, phi2) == pytest.approx((0.6 * 0.3 + 1.0 * 1.0) / 2.0, rel=1e-15)
    # negative value clamps to 0
    assert P.phitphi(x, ab, np.zeros(3)) == 0.0


def test_step_limit_is_the_largest_feasible_step():
    rng = np.random.default_rng(3)
    N, m = 50, 40
    node, normal, _ = rand_rows(rng, N, m)
    x = rng.normal(size=3 * N)
    X = x.reshape(-1, 3)
    c = np.einsum("ij,ij->i", normal, X[node]) - rng.uniform(0.01, 1.0, m)  # strictly feasible
    P = oracle.Projector(problems.rows_to_J(node, normal, N), c)
    d = rng.normal(size=3 * N)
    acg = 1e3
    af = P.step_limit(x, d, acg)
    assert 0 < af <= acg + EPS
    nd = np.einsum("ij,ij->i", normal, d.reshape(-1, 3)[node])
    slack = np.einsum("ij,ij->i", normal, X[node]) - c
    # brute force over rows whose slack shrinks along -d
    cand = [s / v for s, v in zip(slack, nd) if v > EPS and s > 0]
    expect = min([acg + EPS] + cand)
    assert af == pytest.approx(expect, rel=1e-14)
    assert P.is_feasible(x - af * d)
    if af < acg:
        assert not P.is_feasible(x - (af * (1 + 1e-3) + 1e-6) * d)
    assert P.step_limit(x, d) == pytest.approx(min(cand), rel=1e-14)  # alpha_cg = MAX


def test_preconditioner_is_tangent_jacobi():
    rng = np.random.default_rng(4)
    N, m = 30, 20
    A = sp.csr_matrix(random_spd(rng, 3 * N, 50.0) + np.eye(3 * N))
    node, normal, c = rand_rows(rng, N, m)
    P = oracle.Projector(problems.rows_to_J(node, normal, N), c)
    face = (rng.uniform(size=m) < 0.6).astype(np.int8)
    P.set_face(face)
    pc = oracle.Precond(A, P, True)
    g = rng.normal(size=3 * N)
    z = pc.solve(g, P.phi(g)).reshape(-1, 3)
    D = A.diagonal().reshape(-1, 3)
    G = g.reshape(-1, 3)
    act = node[face == 1]
    nact = normal[face == 1]
    # z lies in the tangent plane and D z - g is parallel to n
    np.testing.assert_allclose(np.einsum("ij,ij->i", nact, z[act]), 0, atol=1e-12)
    r = D[act] * z[act] - G[act]
    np.testing.assert_allclose(np.cross(r, nact), 0, atol=1e-12)
    others = np.setdiff1d(np.arange(N), act)
    np.testing.assert_allclose(z[others], G[others] / D[others], rtol=1e-15)
    # no rows at all: the reference silently returns phi (unpreconditioned)
    P0 = oracle.Projector(sp.csr_matrix((0, 3 * N)), np.zeros(0))
    pc0 = oracle.Precond(A, P0, True)
    phi = rng.normal(size=3 * N)
    np.testing.assert_array_equal(pc0.solve(g, phi), phi)


def test_spec_rad_matches_eigvalsh():
    rng = np.random.default_rng(5)
    for n in (9, 30, 60):
        A = random_spd(rng, n, 100.0)
        lam, it, ev = oracle.spec_rad(sp.csr_matrix(A), 1e-12)
        lmax = np.linalg.eigvalsh(A).max()
        assert lam == pytest.approx(lmax, rel=1e-8)
        assert it <= 1001
    # warm start from the returned vector converges immediately
    lam2, it2, _ = oracle.spec_rad(sp.csr_matrix(A), 1e-6, ev=ev)
    assert it2 <= 2 and lam2 == pytest.approx(lmax, rel=1e-8)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("variant", [oracle.MONOTONIC, oracle.CLASSIC])
def test_solve_matches_brute_force_minimiser(seed, variant):
    rng = np.random.default_rng(100 + seed)
    N = 6
    m = int(rng.integers(2, 6))
    A = random_spd(rng, 3 * N, 30.0) + 0.5 * np.eye(3 * N)
    b = rng.normal(size=3 * N) * 3
    node, normal, c = rand_rows(rng, N, m)
    x_bf, f_bf, _ = brute_force_qp(A, b, node, normal, c)
    x0 = rng.normal(size=3 * N)
    # Monotone at tol 1e-8: its acceptance test fy <= fx is decided by rounding
    # once f(x) - f* reaches the objective's ulp, which floors the reachable
    # residual near 1e-8 on these problems (MPRGPSolver.h:288-289).  Classic at
    # 1e-6: its CG steps are never projected (MPRGPSolver.h:182-195) and drift
    # off active rows once the residual nears the rounding floor.
    tol = 1e-8 if variant == oracle.MONOTONIC else 1e-6
    code, x, st, _ = oracle.decoupled_solve(sp.csr_matrix(A), b, problems.rows_to_J(node, normal, N),
                                            c, x0, tol=tol, max_it=10000, variant=variant)
    assert code == 0, st
    lmin = np.linalg.eigvalsh(A)[0]
    np.testing.assert_allclose(x, x_bf, atol=4 * tol / lmin)
    f = 0.5 * x @ A @ x - b @ x
    assert f == pytest.approx(f_bf, rel=1e-6 * tol / 1e-8 if variant else 1e-12, abs=1e-12)


def test_unconstrained_solve_is_cholesky_solve():
    rng = np.random.default_rng(6)
    A = random_spd(rng, 45, 200.0) + np.eye(45)
    b = rng.normal(size=45)
    code, x, st, _ = oracle.decoupled_solve(sp.csr_matrix(A), b, sp.csr_matrix((0, 45)), np.zeros(0),
                                            np.zeros(45), tol=1e-12, max_it=1000)
    assert code == 0
    L = np.linalg.cholesky(A)
    np.testing.assert_allclose(x, np.linalg.solve(L.T, np.linalg.solve(L, b)), rtol=1e-9, atol=1e-11)
    assert st["num_prop"] == 0   # no rows: beta = 0, every point is proportional


def test_reference_test_solver_case():
    # test_case/MPRGPTest.cpp:70-98: n = 5, A = M'M with M = rand + 3I, J = I, random B, c, x
    rng = np.random.default_rng(7)
    n = 5
    M = rng.uniform(-1, 1, (n, n)) + 3 * np.eye(n)
    A = sp.csr_matrix(M.T @ M)
    J = sp.csr_matrix(np.eye(n))
    B, c, x = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    for variant in (oracle.MONOTONIC, oracle.CLASSIC):
        code, y, st, _ = oracle.decoupled_solve(A, B, J, c, x, variant=variant)
        assert code == 0
        assert np.all(y - c >= -1e-9)


def test_max_it_zero_quirk_and_exhaustion():
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    code, x, st, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=0, alpha_bar=1e-3)
    assert code == 0 and st["iterations"] == 0          # residual_out starts at 0
    code, x, st, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=3, alpha_bar=1e-3)
    assert code == -1 and st["iterations"] >= 3


@pytest.mark.parametrize("flat", [False, True])
def test_config1_solution_satisfies_kkt(flat):
    sc = problems.config1_flat() if flat else problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    code, x, st, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-10, max_it=5000)
    assert code == 0, st
    rep = kkt_report(A, b, node, normal, c, x)
    assert rep["infeas"] <= 1e-12
    assert rep["tangential"] <= 1e-9
    assert rep["neg_mult"] <= 1e-9
    assert rep["compl"] <= 1e-9
    frac = st["active"] / sc.n_nodes
    # the 5 degree tilt lands on its lowest corner; the flat drop on its whole
    # bottom face (105 of 525 nodes)
    assert (frac == pytest.approx(0.2)) if flat else (0.0 < frac < 0.05), frac
