"""Pins the CPU oracle's FEM assembly (CPU only): textbook linear-tet
stiffness in Voigt form, rigid-body null spaces, finite differences of the
energy, rotation covariance of the corotated element, and the closed form of
b for a linear body."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from mprgp_amd import problems


def lame(E, nu):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def random_tet(rng):
    X = rng.normal(size=(4, 3)) * 0.05
    if np.linalg.det((X[1:] - X[0]).T) < 0:
        X[[2, 3]] = X[[3, 2]]
    return X


def setup1(X):
    tets, d, vol, mass = oracle.fem_setup(4, np.array([[0, 1, 2, 3]]), X.ravel())
    return d[0], vol[0]


def textbook_linear_K(X, lam, mu):
    """V B^T D B with engineering shear strains (independent of the oracle)."""
    Dm = (X[1:] - X[0]).T
    V = abs(np.linalg.det(Dm)) / 6.0
    Gr = np.linalg.inv(Dm)             # rows: gradients of barycentric coords 1..3
    grads = np.vstack([-Gr.sum(0), Gr])  # [4,3]
    B = np.zeros((6, 12))
    for i in range(4):
        gx, gy, gz = grads[i]
        B[:, 3 * i:3 * i + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz],
                                 [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return V * B.T @ D @ B, V


def test_setup_volumes_dinv_masses():
    rng = np.random.default_rng(0)
    X, T = problems.box_mesh(3, 2, 2, 0.1)
    X = problems.place(X, 7.0, 0.3)
    T2 = T.copy()
    T2[::3, [2, 3]] = T2[::3, [3, 2]]   # flip some orientations
    tets, d, vol, mass = oracle.fem_setup(X.shape[0], T2, X.ravel())
    np.testing.assert_array_equal(tets, T)   # makePositive restored the orientation
    for t in range(tets.shape[0]):
        Dm = (X[tets[t, 1:]] - X[tets[t, 0]]).T
        assert vol[t] == pytest.approx(abs(np.linalg.det(Dm)) / 6, rel=1e-13)
        np.testing.assert_allclose(d[t].reshape(3, 3) @ Dm, np.eye(3), atol=1e-12)
    assert mass.sum() == pytest.approx(vol.sum(), rel=1e-14)
    assert vol.sum() == pytest.approx(3 * 0.1 * 2 * 0.1 * 2 * 0.1, rel=1e-12)


def test_polar_rotation():
    rng = np.random.default_rng(1)
    for _ in range(20):
        Q = problems.random_rotation(rng)
        L = rng.normal(size=(3, 3))
        S = L @ L.T + 0.2 * np.eye(3)
        R = oracle.polar_rotation(Q @ S)
        np.testing.assert_allclose(R, Q, atol=1e-10)
    # inverted element: R is still a rotation and R^T F symmetric
    F = np.diag([1.0, 0.8, -0.3]) @ problems.random_rotation(rng)
    R = oracle.polar_rotation(F)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
    S = R.T @ F
    np.testing.assert_allclose(S, S.T, atol=1e-12)


def test_linear_element_is_textbook():
    rng = np.random.default_rng(2)
    lam, mu = lame(1e5, 0.45)
    for _ in range(5):
        X = random_tet(rng)
        d, V = setup1(X)
        K_tb, V_tb = textbook_linear_K(X, lam, mu)
        assert V == pytest.approx(V_tb, rel=1e-13)
        E, fB, K = oracle.fem_element(oracle.LINEAR, X, d, V, lam, mu)
        np.testing.assert_allclose(K, K_tb, rtol=1e-11, atol=1e-11 * np.abs(K_tb).max())
        # at rest: no force, no energy
        assert abs(E) < 1e-20 and np.abs(fB).max() < 1e-9 * np.abs(K).max()
        # f = -K (x - X) for a linear element
        u = rng.normal(size=12) * 1e-3
        E2, f2, _ = oracle.fem_element(oracle.LINEAR, X.ravel() + u, d, V, lam, mu)
        np.testing.assert_allclose(f2, -K @ u, rtol=1e-9, atol=1e-9 * np.abs(f2).max())
        assert E2 == pytest.approx(0.5 * u @ K @ u, rel=1e-9)


def test_linear_null_space():
    rng = np.random.default_rng(3)
    lam, mu = lame(2e5, 0.3)
    X = random_tet(rng)
    d, V = setup1(X)
    _, _, K = oracle.fem_element(oracle.LINEAR, X, d, V, lam, mu)
    s = np.abs(K).max()
    for k in range(3):
        t = np.zeros(12)
        t[k::3] = 1.0
        assert np.abs(K @ t).max() < 1e-12 * s
        w = np.eye(3)[k]
        rot = np.cross(w, X).ravel()   # infinitesimal rotation
        assert np.abs(K @ rot).max() < 1e-11 * s * np.abs(X).max()


@pytest.mark.parametrize("mtype", [oracle.LINEAR, oracle.COROTATIONAL])
def test_force_is_minus_energy_gradient(mtype):
    rng = np.random.default_rng(4)
    lam, mu = lame(1e5, 0.4)
    X = random_tet(rng)
    d, V = setup1(X)
    Q = problems.random_rotation(rng)
    x = (X @ Q.T + 0.01 * rng.normal(size=(4, 3))).ravel()
    E, fB, K = oracle.fem_element(mtype, x, d, V, lam, mu)
    for _ in range(4):
        u = rng.normal(size=12)
        h = 1e-6
        Ep = oracle.fem_element(mtype, x + h * u, d, V, lam, mu)[0]
        Em = oracle.fem_element(mtype, x - h * u, d, V, lam, mu)[0]
        assert (Ep - Em) / (2 * h) == pytest.approx(-fB @ u, rel=1e-6, abs=1e-9 * np.abs(fB).max())
    if mtype == oracle.LINEAR:  # Hessian by differences of the force
        u = rng.normal(size=12)
        h = 1e-6
        fp = oracle.fem_element(mtype, x + h * u, d, V, lam, mu)[1]
        fm = oracle.fem_element(mtype, x - h * u, d, V, lam, mu)[1]
        np.testing.assert_allclose(-(fp - fm) / (2 * h), K @ u, rtol=1e-6, atol=1e-6 * np.abs(K @ u).max())


def test_corotated_element_covariance():
    rng = np.random.default_rng(5)
    lam, mu = lame(1e5, 0.45)
    X = random_tet(rng)
    d, V = setup1(X)
    # R = I (pure stretch, symmetric F near I) gives the linear element
    Ssym = np.eye(3) + 0.01 * (lambda A: A + A.T)(rng.normal(size=(3, 3)))
    x = X @ Ssym.T
    _, fl, Kl = oracle.fem_element(oracle.LINEAR, x, d, V, lam, mu)
    _, fc, Kc = oracle.fem_element(oracle.COROTATIONAL, x, d, V, lam, mu)
    np.testing.assert_allclose(Kc, Kl, rtol=1e-10, atol=1e-10 * np.abs(Kl).max())
    np.testing.assert_allclose(fc, fl, rtol=1e-8, atol=1e-10 * np.abs(fl).max())
    # rotating the deformed element by Q rotates force and stiffness
    Q = problems.random_rotation(rng)
    y = x @ Q.T
    _, fr, Kr = oracle.fem_element(oracle.COROTATIONAL, y, d, V, lam, mu)
    BQ = np.kron(np.eye(4), Q)
    np.testing.assert_allclose(Kr, BQ @ Kc @ BQ.T, rtol=1e-9, atol=1e-9 * np.abs(Kc).max())
    np.testing.assert_allclose(fr, BQ @ fc, rtol=1e-8, atol=1e-9 * np.abs(fc).max())


def test_build_system_linear_closed_form():
    sc = problems.config1()
    N = sc.n_nodes
    tets, d, vol, mass = oracle.fem_setup(N, sc.tets, sc.X0.ravel())
    A, b = oracle.fem_build_system(N, tets, d, vol, mass, sc.young, sc.poisson, sc.density,
                                   sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, oracle.LINEAR)
    # independent assembly of K from textbook elements
    lam, mu = lame(1e5, 0.45)
    rows, cols, vals = [], [], []
    for t in range(tets.shape[0]):
        Ke, _ = textbook_linear_K(sc.X0[tets[t]], lam, mu)
        dof = (3 * tets[t][:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(dof, 12))
        cols.append(np.tile(dof, 12))
        vals.append(Ke.ravel())
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * N, 3 * N))
    Mv = np.repeat(sc.density * mass, 3)
    A_ref = (sc.h ** 2) * K + sp.diags(Mv)
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    assert abs(A - A_ref).max() <= 1e-10 * abs(A).max()
    assert Mv.sum() == pytest.approx(3 * sc.density * 1.0 * 0.2 * 0.2, rel=1e-12)
    g = np.tile(sc.gravity, N)
    b_ref = Mv * (sc.x_n.ravel() + sc.h * sc.v_n.ravel()) + sc.h ** 2 * (K @ sc.X0.ravel() + Mv * g)
    np.testing.assert_allclose(b, b_ref, rtol=1e-10, atol=1e-10 * np.abs(b_ref).max())
