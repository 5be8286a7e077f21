"""Pins the oracle's time step (CPU only): assembly at a Newton iterate x1,
the implicit-Euler velocity, and the Newton loop of forward.

Checks: for a LINEAR body b does not depend on x1 (Newton on a quadratic
energy); for a COROTATIONAL body b - A x1 - M(x_n + h v_n - x1) - h^2 M g is
h^2 times minus the finite-difference gradient of the total energy at x1; the
velocity is (x - x_n)/h to rounding; a rigid free fall of one step from the
rest shape lands at X0 + h v + h^2 g with velocity v + h g after two Newton
solves (the second one finds the first's solution).  Rayleigh damping: the
damped A and b against the undamped ones, and implicit Euler's discrete
energy identity for a damped free vibration."""
import numpy as np
import pytest

import oracle
from mprgp_amd import problems


def body(nx, ny, nz, E=1e5, nu=0.4, rho=1000.0, a=0.05):
    X, tets = problems.box_mesh(nx, ny, nz, a)
    tets = problems.orient_positive(X, tets)
    N = X.shape[0]
    ot, d, vol, mass = oracle.fem_setup(N, tets, X.ravel())
    T = ot.shape[0]
    return dict(X=X, N=N, tets=ot, d=d, vol=vol, mass=mass, young=np.full(T, E), poisson=np.full(T, nu),
                rho=rho)


def system(b, x_n, v_n, x1, mtype, h=0.01, g=(0.0, -9.8, 0.0)):
    return oracle.fem_build_system(b["N"], b["tets"], b["d"], b["vol"], b["mass"], b["young"], b["poisson"],
                                   b["rho"], x_n, v_n, np.asarray(g, float), h, mtype, x1=x1)


def total_energy(b, x, mtype):
    E, nu = b["young"][0], b["poisson"][0]
    lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    x = x.reshape(-1, 3)
    return sum(oracle.fem_element(mtype, x[t].ravel(), b["d"][k], b["vol"][k], lam, mu)[0]
               for k, t in enumerate(b["tets"]))


def test_linear_rhs_does_not_depend_on_x1():
    sc = problems.config1()
    rng = np.random.default_rng(1)
    b = body(20, 4, 4, E=1e5, nu=0.45)
    x_n = sc.x_n.ravel()
    v_n = sc.v_n.ravel()
    A0, b0 = system(b, x_n, v_n, None, oracle.LINEAR)
    x1 = x_n + 1e-3 * rng.normal(size=x_n.size)
    A1, b1 = system(b, x_n, v_n, x1, oracle.LINEAR)
    assert abs(A1 - A0).max() == 0.0
    np.testing.assert_allclose(b1, b0, rtol=0, atol=1e-11 * np.abs(b0).max())
    # and x1 = None is x1 = x_n
    _, b2 = system(b, x_n, v_n, x_n.copy(), oracle.LINEAR)
    assert np.array_equal(b2, b0)


def test_corotated_rhs_at_x1_is_the_implicit_euler_residual():
    rng = np.random.default_rng(2)
    b = body(2, 1, 1)
    Q = problems.random_rotation(rng)
    X = b["X"]
    x_n = ((X - X.mean(0)) @ Q.T + 0.002 * rng.normal(size=X.shape)).ravel()
    v_n = rng.normal(size=x_n.size)
    x1 = x_n + 0.003 * rng.normal(size=x_n.size)
    h, g = 0.01, np.array([0.0, -9.8, 0.0])
    A, rhs = system(b, x_n, v_n, x1, oracle.COROTATIONAL, h=h, g=g)
    Mv = np.repeat(b["rho"] * b["mass"], 3)
    r = rhs - A @ x1 - Mv * (x_n + h * v_n - x1) - h * h * Mv * np.tile(g, b["N"])
    grad = np.empty(x1.size)
    eps = 1e-7
    for i in range(x1.size):
        e = np.zeros(x1.size)
        e[i] = eps
        grad[i] = (total_energy(b, x1 + e, oracle.COROTATIONAL) -
                   total_energy(b, x1 - e, oracle.COROTATIONAL)) / (2 * eps)
    np.testing.assert_allclose(r, -h * h * grad, rtol=0, atol=1e-5 * np.abs(h * h * grad).max())


def test_velocity_is_the_position_difference():
    rng = np.random.default_rng(3)
    x_n = rng.normal(size=300)
    v_n = rng.normal(size=300)
    x = x_n + 0.01 * rng.normal(size=300)
    v = oracle.implicit_euler_velocity(x_n, v_n, x, 0.01)
    np.testing.assert_allclose(v, (x - x_n) / 0.01, rtol=0, atol=1e-11)


def step(b, x_n, v_n, mtype, planes, **kw):
    return oracle.advance(b["N"], b["tets"], b["d"], b["vol"], b["mass"], b["young"], b["poisson"], b["rho"],
                          x_n, v_n, np.array([0.0, -9.8, 0.0]), 0.01, mtype, planes, **kw)


def test_rigid_free_fall():
    b = body(3, 2, 2)
    X = b["X"].ravel()
    v0 = np.tile([0.3, -1.0, 0.2], b["N"])
    x, v, info = step(b, X, v0, oracle.LINEAR, np.zeros((0, 5)), tol=1e-12, newton_eps=1e-10)
    g = np.tile([0.0, -9.8, 0.0], b["N"])
    np.testing.assert_allclose(x, X + 0.01 * v0 + 1e-4 * g, rtol=0, atol=1e-9)
    np.testing.assert_allclose(v, v0 + 0.01 * g, rtol=0, atol=1e-7)
    assert info["rows"] == 0 and info["newton_iters"] == 2 and info["delta"] == 0.0
    assert info["codes"] == [0, 0]


@pytest.mark.parametrize("mtype", [oracle.LINEAR, oracle.COROTATIONAL])
def test_newton_loop_on_the_ground(mtype):
    sc = problems.config1_flat()
    b = body(20, 4, 4, E=1e5, nu=0.45)
    planes = [[0.0, 1.0, 0.0, 0.0, np.inf]]
    x, v, info = step(b, sc.x_n.ravel(), sc.v_n.ravel(), mtype, planes, tol=1e-8, newton_eps=1e-7,
                      newton_max_it=6)
    assert info["rows"] == b["N"]
    assert all(c == 0 for c in info["codes"])
    if mtype == oracle.LINEAR:
        # the second solve starts at the first one's solution of the same QP
        assert info["newton_iters"] == 2 and info["delta"] == 0.0
    else:
        assert 2 <= info["newton_iters"] <= 6 and info["delta"] < 1e-7
    assert np.all(x.reshape(-1, 3)[:, 1] >= -1e-9)


@pytest.mark.parametrize("mtype", [oracle.LINEAR, oracle.COROTATIONAL])
def test_damped_system_against_the_undamped_one(mtype):
    rng = np.random.default_rng(6)
    b = body(4, 2, 2)
    X = b["X"]
    Q = problems.random_rotation(rng)
    x_n = ((X - X.mean(0)) @ Q.T + 0.002 * rng.normal(size=X.shape)).ravel()
    v_n = rng.normal(size=x_n.size)
    x1 = x_n + 0.003 * rng.normal(size=x_n.size)
    h, ak, am = 0.01, 0.02, 0.01
    g = np.array([0.0, -9.8, 0.0])
    Mv = np.repeat(b["rho"] * b["mass"], 3)
    A0, b0 = system(b, x_n, v_n, x1, mtype, h=h, g=g)
    kw = dict(x1=x1, alpha_k=ak, alpha_m=am)
    A, rhs = oracle.fem_build_system(b["N"], b["tets"], b["d"], b["vol"], b["mass"], b["young"], b["poisson"],
                                     b["rho"], x_n, v_n, g, h, mtype, **kw)
    K = (A0 - oracle_sp_diags(Mv)) / (h * h)
    A_ref = K * (h * h + h * am) + oracle_sp_diags(Mv * (1 + h * ak))
    assert abs(A - A_ref).max() <= 1e-12 * abs(A).max()
    # the damping terms of b: alpha_m K c + M c alpha_k, c = (x_n - x1) h; and A x1 moves with A
    c = (x_n - x1) * h
    b_ref = b0 + am * (K @ c) + Mv * c * ak + (A - A0) @ x1
    np.testing.assert_allclose(rhs, b_ref, rtol=0, atol=1e-11 * np.abs(b0).max())
    # at x1 = x_n the damping leaves b but for A x_n
    A1, b1 = oracle.fem_build_system(b["N"], b["tets"], b["d"], b["vol"], b["mass"], b["young"], b["poisson"],
                                     b["rho"], x_n, v_n, g, h, mtype, alpha_k=ak, alpha_m=am)
    A00, b00 = system(b, x_n, v_n, None, mtype, h=h, g=g)
    np.testing.assert_allclose(b1, b00 + (A1 - A00) @ x_n, rtol=0, atol=1e-11 * np.abs(b00).max())


def oracle_sp_diags(v):
    import scipy.sparse as sp
    return sp.diags(v)


def test_damped_steps_satisfy_the_discrete_energy_identity():
    """A stretched linear bar released with no gravity and no planes.  Implicit
    Euler with Rayleigh damping C = alpha_k M + alpha_m K satisfies, step by
    step, E_{n+1} - E_n = -|dv|_M^2/2 - |du|_K^2/2 - h v_{n+1}' C v_{n+1}
    with E = v'Mv/2 + u'Ku/2, u = x - X0 (exactly, for an exact linear solve)."""
    b = body(4, 1, 1)
    X = b["X"].ravel()
    x = (b["X"] * np.array([1.05, 1.0, 1.0])).ravel()
    v = np.zeros_like(x)
    h, ak, am = 0.001, 0.5, 0.002
    Mv = np.repeat(b["rho"] * b["mass"], 3)
    A0, _ = system(b, X, v, None, oracle.LINEAR, h=h, g=np.zeros(3))
    K = (A0 - oracle_sp_diags(Mv)) / (h * h)

    def energy(x, v):
        u = x - X
        return 0.5 * v @ (Mv * v) + 0.5 * u @ (K @ u)

    e0 = energy(x, v)
    for k in range(10):
        x1, v1, info = oracle.advance(b["N"], b["tets"], b["d"], b["vol"], b["mass"], b["young"], b["poisson"],
                                      b["rho"], x, v, np.zeros(3), h, oracle.LINEAR, np.zeros((0, 5)),
                                      newton_max_it=2, newton_eps=1e-12, tol=1e-12, alpha_k=ak, alpha_m=am)
        dv, du = v1 - v, x1 - x
        loss = 0.5 * dv @ (Mv * dv) + 0.5 * du @ (K @ du) + h * (v1 @ (ak * Mv * v1 + am * (K @ v1)))
        assert energy(x1, v1) - energy(x, v) == pytest.approx(-loss, rel=1e-6, abs=1e-9 * e0), k
        assert loss > 0
        x, v = x1, v1
