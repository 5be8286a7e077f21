"""Mixed-precision form: FP32 inner solves inside FP64 iterative refinement.

Every solve reports its error against the FP64 oracle (printed), and is held
to: KKT residual (FP64, at the returned x) <= tol, solution relative error
<= 1e-6 and objective relative gap <= 1e-9 at tol 1e-8, and the same active
set except nodes within 1e-8 of their plane."""
import numpy as np
import pytest

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import kkt_report, scene_qp

pytestmark = pytest.mark.gpu


def slack(x, node, normal, c):
    return np.einsum("ij,ij->i", normal, x.reshape(-1, 3)[node]) - c


def run_pair(ctx, sc, tol):
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=tol, max_it=50000, alpha_bar=ab)
    Ad = ctx.matrix_csr(A)
    q = Ad.qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    rc, sm = q.solve(tol=tol, max_it=50000, alpha_bar=ab, precision=M.MIXED)
    xm = q.get_x()
    q.reset_x()
    rc64, s64 = q.solve(tol=tol, max_it=50000, alpha_bar=ab)
    f = lambda x: 0.5 * x @ (A @ x) - b @ x
    err = np.linalg.norm(xm - xr) / np.linalg.norm(xr)
    gap = abs(f(xm) - f(xr)) / abs(f(xr))
    print(f"{sc.name}: mixed rel err {err:.2e}, objective gap {gap:.2e}, refinements {sm['refinements']}, "
          f"iterations mixed {sm['iterations']} / fp64 gpu {s64['iterations']} / oracle {sr['iterations']}")
    return A, b, node, normal, c, xm, xr, sm, sr, err, gap


@pytest.mark.parametrize("which", ["config1", "config1_flat", "config3"])
def test_mixed_against_fp64_oracle(ctx, which):
    sc = getattr(problems, which)()
    A, b, node, normal, c, xm, xr, sm, sr, err, gap = run_pair(ctx, sc, 1e-8)
    assert sm["code"] == 0, sm
    assert sm["residual"] <= 1e-8
    rep = kkt_report(A, b, node, normal, c, xm)
    assert rep["infeas"] <= 1e-12
    assert err <= 1e-6 and gap <= 1e-9
    sg, srr = slack(xm, node, normal, c), slack(xr, node, normal, c)
    differ = (np.abs(sg) < 1e-9) != (np.abs(srr) < 1e-9)
    near = (np.abs(sg) < 1e-8) | (np.abs(srr) < 1e-8)
    assert not np.any(differ & ~near)


def test_mixed_batch_equals_alone(ctx):
    scenes = [problems.config1(seed=s) for s in (21, 22, 23)]
    qs, alone = [], []
    for sc in scenes:
        A, b, node, normal, c, x0 = scene_qp(sc)
        q = ctx.matrix_csr(A).qp()
        q.set_rows(node, normal, c)
        q.set_rhs(b)
        q.set_x(x0)
        ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
        rc, st = q.solve(tol=1e-7, max_it=5000, alpha_bar=ab, precision=M.MIXED)
        alone.append((q.get_x(), st))
        qs.append((q, ab))
    for q, _ in qs:
        q.reset_x()
    rc, stats = M.solve_batch(ctx, [q for q, _ in qs],
                              per_qp=[M.params(tol=1e-7, max_it=5000, alpha_bar=ab, precision=M.MIXED)
                                      for _, ab in qs])
    for (q, _), (xa, sa), sb in zip(qs, alone, stats):
        np.testing.assert_array_equal(q.get_x(), xa)
        assert sb["iterations"] == sa["iterations"] and sb["refinements"] == sa["refinements"]


def test_full_batch_of_another_seed(ctx):
    """The benchmark's workload shape (256 beams, 5k-60k DoF each) from a
    seed the benchmark does not use: every QP converges in both forms, the
    FP64 KKT residual at the returned x is <= tol, and the mixed solutions
    stay within 1e-5 of the FP64 ones (tol 1e-6 on the KKT residual bounds
    the error by tol / lambda_min(A) ~ 1e-4 absolute on coordinates of
    order 1)."""
    from mprgp_amd import workloads
    scenes = problems.batch_scenes(256, seed=12, tol=1e-6, max_it=4000)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=1e-6, max_it=0, power_tol=1e-6))
    ab = [s["alpha_bar"] for s in st]
    xs = {}
    for prec in (M.FP64, M.MIXED):
        per = [M.params(tol=1e-6, max_it=4000, alpha_bar=a, project_first=1, precision=prec) for a in ab]
        for q in qps:
            q.qp.reset_x()
        rc, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
        assert rc == 0 and all(s["code"] == 0 for s in st), prec
        assert max(s["residual"] for s in st) <= 1e-6, prec
        xs[prec] = [q.qp.get_x() for q in qps]
        assert all(np.all(np.isfinite(x)) for x in xs[prec])
    rel = [np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(xs[M.MIXED], xs[M.FP64])]
    print(f"mixed vs fp64: max {max(rel):.2e}, median {np.median(rel):.2e}")
    assert max(rel) <= 1e-5


def test_streamed_and_cached_variants_agree_bitwise(ctx, monkeypatch):
    """The step kernel's streamed variant (matrix values and column indices
    loaded nontemporal, launched while the live QPs' matrices exceed the
    last-level cache) and its cached variant give the same bits: one batch
    solved with the streaming threshold at 0 bytes (every launch above the
    tail size streams) and at 1e18 bytes (none does), in both forms."""
    from mprgp_amd import workloads
    scenes = problems.batch_scenes(16, seed=31, tol=1e-6, max_it=4000)
    qps = [workloads.DeviceQP(ctx, sc) for sc in scenes]
    rc, st = M.solve_batch(ctx, [q.qp for q in qps], M.params(tol=1e-6, max_it=0, power_tol=1e-6))
    ab = [s["alpha_bar"] for s in st]
    for prec in (M.FP64, M.MIXED):
        out = {}
        for thr in ("0", "1e18"):
            monkeypatch.setenv("MPRGP_STREAM_BYTES", thr)
            for q in qps:
                q.qp.reset_x()
            per = [M.params(tol=1e-6, max_it=4000, alpha_bar=a, project_first=1, precision=prec, tiles=1)
                   for a in ab]
            rc, st = M.solve_batch(ctx, [q.qp for q in qps], per_qp=per)
            assert all(s["code"] == 0 for s in st), (prec, thr)
            out[thr] = ([q.qp.get_x() for q in qps], [s["iterations"] for s in st])
        for a, b in zip(out["0"][0], out["1e18"][0]):
            np.testing.assert_array_equal(a, b)
        assert out["0"][1] == out["1e18"][1]


@pytest.mark.parametrize("which", ["config1", "config3"])
def test_error_bound_of_every_solve(ctx, which):
    """mprgp_stats.error_bound = residual / lambda_min_lower, with the bound
    lambda_min(A) >= min lumped mass that the device FEM assembly records:
    the mixed solve's distance from the oracle's solution (to 1e-11) is
    within its bound plus the oracle's, and a matrix without a known bound
    reports -1."""
    from mprgp_amd import workloads
    sc = getattr(problems, which)()
    dq = workloads.DeviceQP(ctx, sc)
    tets, d, vol, mass = oracle.fem_setup(sc.n_nodes, sc.tets, sc.X0)
    A, b = oracle.fem_build_system(sc.n_nodes, tets, d, vol, mass, sc.young, sc.poisson, sc.density,
                                   sc.x_n.ravel(), sc.v_n.ravel(), sc.gravity, sc.h, sc.material)
    lam_lo = float(np.min(mass)) * sc.density
    node, normal, c = sc.rows()
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, sc.x_n.ravel(), tol=1e-11, max_it=50000,
                                             alpha_bar=ab, project_first=True)
    assert code == 0
    # the bound really is below A's smallest eigenvalue
    assert lam_lo <= np.linalg.eigvalsh(A.toarray())[0] * (1 + 1e-12)
    for tol in (1e-4, 1e-6, 1e-8):
        dq.qp.reset_x()
        rc, st = dq.qp.solve(tol=tol, max_it=50000, alpha_bar=ab, precision=M.MIXED, project_first=1)
        assert rc == 0
        err = float(np.linalg.norm(dq.qp.get_x() - xr))
        print(f"{which} tol {tol:g}: |x - x_oracle| {err:.2e}, error_bound {st['error_bound']:.2e} "
              f"(residual {st['residual']:.2e} / lambda_min >= {lam_lo:.3e})")
        assert st["error_bound"] == pytest.approx(st["residual"] / lam_lo, rel=1e-12)
        assert err <= st["error_bound"] + 1e-11 / lam_lo
    Ad = ctx.matrix_csr(A)  # uploaded as plain arrays: no bound known
    q = Ad.qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(sc.x_n.ravel())
    rc, st = q.solve(tol=1e-6, max_it=50000, alpha_bar=ab, precision=M.MIXED, project_first=1)
    assert rc == 0 and st["error_bound"] == -1.0
    Ad.set_lambda_min(lam_lo)
    q.reset_x()
    rc, st = q.solve(tol=1e-6, max_it=50000, alpha_bar=ab, precision=M.MIXED, project_first=1)
    assert st["error_bound"] == pytest.approx(st["residual"] / lam_lo, rel=1e-12)


def chain_qp(n, k=40.0, load=1.0):
    """A chain of n nodes: A = (1 + k) I on the diagonal blocks, -k/2 I to
    each neighbour (eigenvalues in (1, 1 + 2k)); one row y >= 0 on every 7th
    node; b pulls every node down by `load`, so each row carries about 7 load."""
    import scipy.sparse as sp
    T = sp.diags([np.full(n - 1, -0.5 * k), np.full(n, 1.0 + k), np.full(n - 1, -0.5 * k)], [-1, 0, 1])
    A = sp.kron(T, sp.identity(3)).tocsr()
    i = np.arange(n)
    b = np.stack([0.01 * np.sin(0.1 * i), -load * np.ones(n), 0.01 * np.cos(0.1 * i)], 1).ravel()
    x0 = np.zeros(3 * n)
    x0[1::3] = 0.05
    node = i[i % 7 == 0]
    normal = np.tile([0.0, 1.0, 0.0], (node.size, 1))
    return A, b, x0, node, normal, np.zeros(node.size)


def test_stalled_refinement_finishes_in_fp64(ctx):
    """Contact forces of about 7 per row put the FP32 inner solve's floor
    (eps32 |lambda| per active row) above tol = 1e-6: every refinement sweep
    runs and the residual stays above tol, so the FP64 form finishes from that
    x.  The result converges (FP64 KKT residual <= tol) and lies within the
    two solves' error bounds of the plain FP64 solve."""
    A, b, x0, node, normal, c = chain_qp(12000)
    ab = 2.0 / (1.0 + 2 * 40.0)  # Gershgorin bound on the largest eigenvalue
    m = ctx.matrix_csr(A)
    m.set_lambda_min(1.0)       # Gershgorin bound on the smallest
    q = m.qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    out = {}
    for prec in (M.FP64, M.MIXED):
        q.reset_x()
        rc, st = q.solve(M.params(tol=1e-6, max_it=20000, alpha_bar=ab, precision=prec))
        print(prec, {k: st[k] for k in ("code", "iterations", "refinements", "fp64_finish", "residual")})
        assert rc == 0 and st["code"] == 0 and st["residual"] <= 1e-6
        out[prec] = (q.get_x(), st)
    xm, sm = out[M.MIXED]
    xf, sf = out[M.FP64]
    assert sm["fp64_finish"] == 1 and sm["refinements"] == 8 and sf["fp64_finish"] == 0
    assert np.linalg.norm(xm - xf) <= sm["error_bound"] + sf["error_bound"]
    # a QP inside the mixed form's range never takes the FP64 finish
    sc = problems.config1()
    Ab, bb, nb, nrm, cb, xb = scene_qp(sc)
    qb = ctx.matrix_csr(Ab).qp()
    qb.set_rows(nb, nrm, cb)
    qb.set_rhs(bb)
    qb.set_x(xb)
    rc, st = qb.solve(M.params(tol=1e-6, max_it=20000, precision=M.MIXED, project_first=1))
    assert rc == 0 and st["fp64_finish"] == 0
