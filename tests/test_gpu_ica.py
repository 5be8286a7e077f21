"""GPU parity of the ICA comparator (mprgp_ica_solve: ICASolver,
ICASolver.h:287-463) with the oracle's restatement.

The device runs the oracle's arithmetic in the oracle's order (the sweep in
the user's node order by dependency levels, every row's terms singly rounded
in column order; the projected Gauss-Seidel on the host), so the iterates
agree to rounding: the same iteration count, and x and lambda within 1e-10
relative.  The converged result is also checked against what fixes it: the
enumerated minimiser on tiny QPs, the MPRGP solution on the beams."""
import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import scene_qp
from tests.test_gpu_general import beam_with_contacts, qp_of
from tests.test_oracle_general import brute_qp, random_coupled
from tests.test_oracle_ica import block_dominant_spd

pytestmark = pytest.mark.gpu


def both(ctx, A, b, J, c, x0, **kw):
    code, xr, lr, sr = oracle.ica_solve(A, b, J, c, x0, **kw)
    q = qp_of(ctx, A, b, x0)
    rc, sg, lg, sweeps = M.ica_solve(q, J, c, **kw)
    xg = q.get_x()
    print(f"oracle {sr['iterations']} it ({sr['gs_iterations']} pgs, res {sr['residual']:.3e}), "
          f"gpu {sg['iterations']} it ({sweeps} pgs, res {sg['residual']:.3e}, {sg['elapsed_ms']:.1f} ms)",
          flush=True)
    assert (rc == 0) == (code == 0) and sg["code"] == code
    assert sg["iterations"] == sr["iterations"] and sweeps == sr["gs_iterations"]
    assert np.linalg.norm(xg - xr) <= 1e-10 * np.linalg.norm(xr)
    assert np.linalg.norm(lg - lr) <= 1e-10 * max(1.0, np.linalg.norm(lr))
    return xg, lg, sg


@pytest.mark.parametrize("seed", range(4))
def test_tiny_coupled_qps(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    n_nodes = 6
    A = block_dominant_spd(rng, n_nodes)
    J = random_coupled(rng, n_nodes=n_nodes, m=4)
    x0 = rng.normal(size=3 * n_nodes)
    b = A @ x0 + rng.normal(size=3 * n_nodes)
    c = J @ x0 - 0.1 * rng.random(4)
    x, lam, st = both(ctx, sp.csr_matrix(A), b, J, c, x0, tol=1e-11, max_it=20000, gs_max_it=200, gs_tol=1e-14)
    assert st["code"] == 0
    np.testing.assert_allclose(x, brute_qp(A, b, J, c), atol=1e-8)
    assert np.all(lam >= 0.0)


def test_no_rows_is_block_gauss_seidel(ctx):
    rng = np.random.default_rng(3)
    A = block_dominant_spd(rng, 40)
    b = rng.normal(size=120)
    x, _, st = both(ctx, sp.csr_matrix(A), b, sp.csr_matrix((0, 120)), np.zeros(0), np.zeros(120), tol=1e-12,
                    max_it=10000)
    assert st["code"] == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-11)


@pytest.mark.parametrize("which,tol", [("config1", 1e-8), ("config1_flat", 1e-8), ("config2", 1e-6)])
def test_beams_with_plane_rows(ctx, which, tol):
    """Per-node plane rows (B diagonal); ICA reaches MPRGP's minimiser."""
    sc = getattr(problems, which)()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    x, lam, st = both(ctx, A, b, J, c, x0, tol=tol, max_it=5000)
    assert st["code"] == 0
    q = qp_of(ctx, A, b, x0)
    q.set_rows(node, normal, c)
    rc, sm = q.solve(tol=tol * 1e-2, max_it=20000, project_first=1)
    xm = q.get_x()
    assert rc == 0
    # both within their tolerances of the same minimiser (lambda_min(A) >= the lumped mass)
    assert np.linalg.norm(x - xm) <= 1e-6 * np.linalg.norm(xm)


def test_beam_with_self_contact_rows(ctx):
    sc, A, b, J, c, x0, m_ground = beam_with_contacts()
    x, lam, st = both(ctx, A, b, J, c, x0, tol=1e-8, max_it=5000)
    assert st["code"] == 0
    s = J @ x - c
    assert s.min() >= -1e-7 and np.all(lam >= 0.0)
    assert np.any(lam[m_ground:] > 0.0), "no coupled row carries force"


def test_ica_errors(ctx):
    rng = np.random.default_rng(0)
    A = block_dominant_spd(rng, 4)
    q = qp_of(ctx, sp.csr_matrix(A), np.ones(12), np.zeros(12))
    J = sp.csr_matrix(np.eye(12)[:2])
    with pytest.raises(M.MprgpError) as e:
        M.ica_solve(q, J, [0.0, 0.0], tol=float("nan"))
    assert e.value.code == -2
    bad = sp.csr_matrix((np.ones(1), (np.zeros(1, int), np.array([40]))), shape=(1, 48))
    with pytest.raises(M.MprgpError) as e:
        M.ica_solve(q, bad, [0.0])
    assert e.value.code == -2 and "out of range" in M.last_error()
    rc, st, lam, sweeps = M.ica_solve(q, J, [0.0, 0.0], tol=0.0, max_it=3)
    assert rc != 0 and st["code"] == -1 and st["iterations"] == 3
