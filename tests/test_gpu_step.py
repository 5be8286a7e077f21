"""GPU parity of the time step (mprgp_fem_advance, begin_step/end_step) and
of the assembly at a Newton iterate (mprgp_fem_assemble_at), with and without
Rayleigh damping, against the oracle's.

Bars: assembly A x and b within 1e-12 (relative to the largest entry, as the
x1 = x_n assembly tests); a step's positions within 1e-8 relative (the FP64
solve bar), velocities within the same error divided by h, the same row
count and Newton iteration count; split and whole one-solve steps bitwise
equal; the rigid free fall against its closed form."""
import numpy as np
import pytest

import mprgp_amd as M
import oracle
from mprgp_amd import _lib, problems

pytestmark = pytest.mark.gpu


def oracle_body(sc):
    tets, d, vol, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
    return tets, d, vol, mass


def deformed(sc, seed, scale=0.002):
    rng = np.random.default_rng(seed)
    Q = problems.random_rotation(rng)
    c = sc.X0.mean(0)
    return ((sc.X0 - c) @ Q.T + c + scale * rng.normal(size=sc.X0.shape)).ravel()


@pytest.mark.parametrize("damping", [(0.0, 0.0), (0.02, 0.01)])
@pytest.mark.parametrize("mtype", [M.LINEAR, M.COROTATIONAL])
def test_assembly_at_newton_iterate_matches_oracle(ctx, mtype, damping):
    sc = problems.config1_flat()
    tets, d, vol, mass = oracle_body(sc)
    x_n = deformed(sc, 4)
    x1 = x_n + 0.003 * np.random.default_rng(5).normal(size=x_n.size)
    ak, am = damping
    A, b = oracle.fem_build_system(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson, sc.density, x_n,
                                   sc.v_n.ravel(), sc.gravity, sc.h, mtype, x1=x1, alpha_k=ak, alpha_m=am)
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    fem.set_damping(ak, am)
    q = fem.matrix.qp()
    fem.assemble_at(q, x_n, sc.v_n.ravel(), x1, sc.gravity, sc.h, mtype)
    rng = np.random.default_rng(0)
    for _ in range(2):
        x = rng.normal(size=A.shape[0])
        scale = abs(A) @ np.abs(x)
        assert np.all(np.abs(fem.matrix.spmv(x) - A @ x) <= 1e-12 * scale)
    bg = q.get_rhs()
    assert np.abs(bg - b).max() <= 1e-12 * np.abs(b).max(), np.abs(bg - b).max() / np.abs(b).max()


def make(ctx, sc):
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    q = fem.matrix.qp()
    fem.set_state(sc.x_n.ravel(), sc.v_n.ravel())
    return fem, q


@pytest.mark.parametrize("mtype,damping", [(M.LINEAR, (0.0, 0.0)), (M.COROTATIONAL, (0.0, 0.0)),
                                           (M.COROTATIONAL, (0.02, 0.01))])
def test_steps_match_oracle(ctx, mtype, damping):
    sc = problems.config1()
    tets, d, vol, mass = oracle_body(sc)
    planes = problems.plane_array(sc.planes)
    fem, q = make(ctx, sc)
    ak, am = damping
    fem.set_damping(ak, am)
    sp = fem.step_params(h=sc.h, gravity=sc.gravity, material=mtype, planes=planes, newton_max_it=4,
                         newton_eps=1e-7)
    x_r, v_r = sc.x_n.ravel(), sc.v_n.ravel()
    alpha_bar = 0.1  # fixed step length on both sides (below 2/|A| for this body)
    for k in range(3):
        rc, st = fem.advance(q, sp, tol=1e-8, max_it=1000, alpha_bar=alpha_bar)
        x_r, v_r, info = oracle.advance(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson, sc.density,
                                        x_r, v_r, sc.gravity, sc.h, mtype, planes, newton_max_it=4,
                                        newton_eps=1e-7, tol=1e-8, max_it=1000, alpha_bar=alpha_bar,
                                        alpha_k=ak, alpha_m=am)
        x_g, v_g = fem.get_state()
        assert rc == 0 and all(c == 0 for c in info["codes"]), (k, st, info["codes"])
        assert st["rows"] == info["rows"], k
        assert st["newton_iters"] == info["newton_iters"], (k, st, info)
        ex = np.linalg.norm(x_g - x_r)
        assert ex <= 1e-8 * np.linalg.norm(x_r), (k, ex)
        assert np.linalg.norm(v_g - v_r) <= 2e-8 * np.linalg.norm(x_r) / sc.h, k
        # keep the two trajectories on the same states
        fem.set_state(x_r, v_r)


def test_split_step_equals_advance(ctx):
    sc = problems.config1_flat()
    planes = problems.plane_array(sc.planes)
    fa, qa = make(ctx, sc)
    fb, qb = make(ctx, sc)
    for k in range(3):
        mat = M.COROTATIONAL if k == 0 else (M.COROTATIONAL | M.RHS_ONLY)
        sp_a = fa.step_params(h=sc.h, gravity=sc.gravity, material=mat, planes=planes, newton_max_it=1)
        sp_b = fb.step_params(h=sc.h, gravity=sc.gravity, material=mat, planes=planes, newton_max_it=1)
        p = M.params(tol=1e-6, max_it=1000, project_first=1, use_ev=1, power_tol=1e-6)
        rc, st = fa.advance(qa, sp_a, p)
        rows = fb.begin_step(qb, sp_b)
        rc2, st2 = qb.solve(p)
        delta = fb.end_step(qb, sp_b)
        assert rc == rc2 == 0 and rows == st["rows"]
        assert delta == st["delta"]
        xa, va = fa.get_state()
        xb, vb = fb.get_state()
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), k


def test_rigid_free_fall(ctx):
    X, tets = problems.box_mesh(3, 2, 2, 0.05)
    tets = problems.orient_positive(X, tets)
    N = X.shape[0]
    fem = M.Fem(ctx, X, tets, 1e5, 0.4, 1000.0)
    q = fem.matrix.qp()
    v0 = np.tile([0.3, -1.0, 0.2], N)
    fem.set_state(X.ravel(), v0)
    sp = fem.step_params(h=0.01, gravity=(0.0, -9.8, 0.0), material=M.LINEAR, planes=np.zeros((0, 5)),
                         newton_max_it=10, newton_eps=1e-10)
    rc, st = fem.advance(q, sp, tol=1e-11, max_it=1000)
    x, v = fem.get_state()
    g = np.tile([0.0, -9.8, 0.0], N)
    # the second solve starts at the first one's solution (g recomputed as A x - b)
    assert rc == 0 and st["rows"] == 0 and st["newton_iters"] == 2 and st["delta"] < 1e-10
    np.testing.assert_allclose(x, X.ravel() + 0.01 * v0 + 1e-4 * g, rtol=0, atol=1e-9)
    np.testing.assert_allclose(v, v0 + 0.01 * g, rtol=0, atol=1e-7)


def test_step_errors(ctx):
    sc = problems.config1()
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    q = fem.matrix.qp()
    sp = fem.step_params(planes=problems.plane_array(sc.planes))
    with pytest.raises(M.MprgpError) as e:
        fem.advance(q, sp)
    assert e.value.code == _lib.ESTATE
    fem.set_state(sc.x_n.ravel(), sc.v_n.ravel())
    with pytest.raises(M.MprgpError) as e:
        fem.end_step(q, sp)
    assert e.value.code == _lib.ESTATE
    bad = fem.step_params(planes=problems.plane_array(sc.planes), newton_max_it=0)
    with pytest.raises(M.MprgpError) as e:
        fem.advance(q, bad)
    assert e.value.code == _lib.EBADINPUT
    for ak, am in ((-0.1, 0.0), (0.0, float("nan"))):
        with pytest.raises(M.MprgpError) as e:
            fem.set_damping(ak, am)
        assert e.value.code == _lib.EBADINPUT
