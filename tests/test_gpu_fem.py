"""GPU parity of the device FEM assembly (A = M + h^2 K, b) against the
oracle's assembly, for LINEAR and COROTATIONAL tets, and of the solve that
consumes it."""
import numpy as np
import pytest

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


def oracle_system(sc, x_n, mtype):
    tets, d, vol, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
    return oracle.fem_build_system(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson, sc.density,
                                   x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, mtype)


def deformed(sc, seed):
    rng = np.random.default_rng(seed)
    Q = problems.random_rotation(rng)
    c = sc.X0.mean(0)
    return (sc.X0 - c) @ Q.T + c + 0.002 * rng.normal(size=sc.X0.shape)


def check_system(ctx, sc, x_n, mtype):
    A, b = oracle_system(sc, x_n, mtype)
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    q = fem.matrix.qp()
    fem.assemble(q, x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, mtype)
    rng = np.random.default_rng(0)
    absA = abs(A)
    for _ in range(3):
        x = rng.normal(size=A.shape[0])
        y = fem.matrix.spmv(x)
        scale = absA @ np.abs(x)
        assert np.all(np.abs(y - A @ x) <= 1e-12 * scale), np.max(np.abs(y - A @ x) / scale)
    bg = q.get_rhs()
    scale_b = np.abs(b).max()
    assert np.abs(bg - b).max() <= 1e-12 * scale_b, np.abs(bg - b).max() / scale_b
    return fem, q, A, b


@pytest.mark.parametrize("mtype", [M.LINEAR, M.COROTATIONAL])
def test_assembly_matches_oracle_config1(ctx, mtype):
    sc = problems.config1()
    check_system(ctx, sc, sc.x_n, mtype)


def test_corotated_assembly_of_rotated_body(ctx):
    sc = problems.config1_flat()
    check_system(ctx, sc, deformed(sc, 1), M.COROTATIONAL)


def test_heterogeneous_material_assembly(ctx):
    sc = problems.config3()
    check_system(ctx, sc, deformed(sc, 2), M.COROTATIONAL)


def test_assembled_system_solves_like_oracle(ctx):
    sc = problems.config1_flat()
    fem, q, A, b = check_system(ctx, sc, sc.x_n, M.LINEAR)
    node, normal, c = sc.rows()
    x0 = problems.project_rows(sc.x_n.ravel(), node, normal, c)
    q.set_rows(node, normal, c)
    q.set_x(x0)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, problems.rows_to_J(node, normal, sc.n_nodes), c, x0,
                                             tol=1e-9, max_it=5000, alpha_bar=ab)
    rc, sg = q.solve(tol=1e-9, max_it=5000, alpha_bar=ab)
    assert rc == 0 == code
    rel = np.linalg.norm(q.get_x() - xr) / np.linalg.norm(xr)
    assert rel <= 1e-8, rel
    assert sg["active"] == sr["active"]


def test_assembly_config2_size(ctx):
    sc = problems.config2()
    check_system(ctx, sc, sc.x_n, M.LINEAR)


@pytest.mark.parametrize("mtype", [M.LINEAR, M.COROTATIONAL])
def test_rhs_only_is_the_kept_linearisation(ctx, mtype):
    """RHS-only assembly at x' keeps A (and, for COROT, the element rotations)
    of the full assembly at x0.  The element energy is then quadratic in x with
    exactly the kept stiffness, so f(x') = f(x0) - K (x' - x0) and
    b(x') = b(x0) + M (x' - x0) with the lumped mass M: a closed form that a
    wrong rotation, a stale force or a missing A x' term all break."""
    sc = problems.config1_flat()
    x0 = deformed(sc, 3)
    fem = M.Fem(ctx, sc.X0, sc.tets, sc.young, sc.poisson, sc.density)
    q = fem.matrix.qp()
    if mtype == M.COROTATIONAL:
        with pytest.raises(M.MprgpError):  # no rotations kept yet
            fem.assemble(q, x0.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, mtype | M.RHS_ONLY)
    fem.assemble(q, x0.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, mtype)
    b0 = q.get_rhs()
    rng = np.random.default_rng(4)
    x1 = x0 + 0.01 * rng.normal(size=x0.shape) + np.array([0.0, -0.05, 0.02])
    probe = rng.normal(size=3 * sc.n_nodes)
    Ax = fem.matrix.spmv(probe)
    fem.assemble(q, x1.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, mtype | M.RHS_ONLY)
    b1 = q.get_rhs()
    np.testing.assert_array_equal(fem.matrix.spmv(probe), Ax)  # matrix kept
    _, _, _, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
    m3 = np.repeat(mass * sc.density, 3)
    expect = b0 + m3 * (x1 - x0).ravel()
    err = np.abs(b1 - expect).max() / np.abs(b1).max()
    assert err <= 1e-12, err


def test_assembled_matrix_downloads_as_the_oracle_system(ctx):
    """mprgp_matrix_get_bsr3 returns the device-assembled matrix in the
    oracle's pattern and values (entrywise within 1e-12 of the row scale),
    and a matrix uploaded from what it returns multiplies bit for bit as the
    original; with mprgp_write_qp that is the simulator's per-frame QP dump
    (saveQP, MprgpFemSolver.h:129-133)."""
    sc = problems.config1()
    fem, q, A, b = check_system(ctx, sc, deformed(sc, 4), M.COROTATIONAL)
    Ad = fem.matrix.to_csr()
    Ad.sort_indices()
    Ar = A.tocsr()
    Ar.sort_indices()
    assert np.array_equal(Ad.indptr, Ar.indptr) and np.array_equal(Ad.indices, Ar.indices)
    row_scale = np.repeat(np.abs(Ar).max(axis=1).toarray().ravel(), np.diff(Ar.indptr))
    assert np.all(np.abs(Ad.data - Ar.data) <= 1e-12 * row_scale)
    rp, ci, bl = fem.matrix.get_bsr3()
    again = ctx.matrix_bsr3(sc.n_nodes, rp, ci, bl.reshape(-1))
    x = np.random.default_rng(5).normal(size=3 * sc.n_nodes)
    np.testing.assert_array_equal(again.spmv(x), fem.matrix.spmv(x))
