"""GPU parity of the projector and preconditioner entry points (the step
kernel's per-node code applied once) against the oracle's DecoupledConProjector
and DiagonalDecouplePreconSolver restatements.

Bars: per-node results within 1e-14 of the entry scale (FMA and summation
order differ), reductions within 1e-12 relative, faces, active counts and
feasibility identical (inputs are built with every row either on its plane to
~1e-16 or at least 1e-3 away)."""
import numpy as np
import pytest

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(ctx):
    sc = problems.config1()
    A, b, *_ = scene_qp(sc)
    N = sc.n_nodes
    rng = np.random.default_rng(21)
    m = 200
    node = rng.choice(N, size=m, replace=False).astype(np.int32)
    normal = rng.normal(size=(m, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    x = rng.normal(size=3 * N)
    nx = np.einsum("ij,ij->i", normal, x.reshape(-1, 3)[node])
    # rows 0..79 on their plane, 80..139 strictly feasible, the rest violated
    off = np.concatenate([np.zeros(80), -rng.uniform(1e-3, 1.0, 60), rng.uniform(1e-3, 1.0, 60)])
    c = nx + off
    q = ctx.matrix_csr(A).qp()
    q.set_rows(node, normal, c)
    P = oracle.Projector(problems.rows_to_J(node, normal, N), c)
    return dict(A=A, N=N, q=q, P=P, x=x, rng=rng, node=node, normal=normal, c=c)


def close(a, b, tol=1e-14):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, np.abs(a - b).max() / scale


def test_project_and_feasibility(setup):
    q, P, x = setup["q"], setup["P"], setup["x"]
    y = q.project(x)
    close(y, P.project(x))
    assert not q.is_feasible(x) and not P.is_feasible(x)
    assert q.is_feasible(y) and P.is_feasible(y)
    # projection is idempotent on the device too
    close(q.project(y), y)


def test_face_phi_beta(setup):
    q, P, x, rng = setup["q"], setup["P"], setup["x"], setup["rng"]
    k = q.decide_face(x)
    P.decide_face(x)
    assert k == P.num_active() == 80
    assert list(P.face()) == [1] * 80 + [0] * 120
    face_o = np.zeros(setup["N"], dtype=np.uint8)
    face_o[setup["node"][:80]] = 1
    assert np.array_equal(q.get_face(), face_o)
    g = rng.normal(size=x.size)
    phi = q.phi(g)
    close(phi, P.phi(g))
    close(q.beta(g, phi), P.beta(g, phi))


def test_phitphi_and_step_limit(setup):
    q, P, x, rng = setup["q"], setup["P"], setup["x"], setup["rng"]
    y = q.project(x)
    phi = rng.normal(size=x.size)
    for ab in (0.05, 0.5):
        a, r = q.phitphi(y, ab, phi), P.phitphi(y, ab, phi)
        assert a == pytest.approx(r, rel=1e-12, abs=1e-300)
    # every row strictly feasible by at least 1e-3 (a row on its plane would
    # make the n.x > c test a matter of rounding)
    ys = y.reshape(-1, 3).copy()
    ys[setup["node"]] += setup["normal"] * rng.uniform(1e-3, 0.1, setup["node"].size)[:, None]
    ys = ys.ravel()
    d = rng.normal(size=x.size)
    for acg in (1e-3, 10.0, np.finfo(np.float64).max):
        a, r = q.step_limit(ys, d, acg), P.step_limit(ys, d, acg)
        assert a == pytest.approx(r, rel=1e-12), acg
    assert P.step_limit(ys, d) < np.finfo(np.float64).max  # some row limits the step


def test_precond(setup):
    q, P, x, rng, A = setup["q"], setup["P"], setup["x"], setup["rng"], setup["A"]
    q.decide_face(x)
    P.decide_face(x)
    g = rng.normal(size=x.size)
    phi = P.phi(g)
    close(q.precond(g, phi), oracle.Precond(A, P).solve(g, phi))


def test_no_rows(ctx, setup):
    q = ctx.matrix_csr(setup["A"]).qp()
    rng = setup["rng"]
    x, g, phi = (rng.normal(size=3 * setup["N"]) for _ in range(3))
    assert np.array_equal(q.project(x), x) and q.is_feasible(x) and q.decide_face(x) == 0
    assert np.array_equal(q.precond(g, phi), phi)  # the reference's unpreconditioned case
    assert q.step_limit(x, g, 1.0) == 1.0 + 1e-9
    with pytest.raises(M.MprgpError):
        q.phitphi(x, 0.0, phi)
