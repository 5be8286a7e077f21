"""GPU parity: the HIP solve through the C ABI against the CPU oracle.

Bars (FP64 form): solution relative error <= 1e-8, the same active set except
nodes within 1e-10 of their plane, objective relative error <= 1e-10; SpMV
relative error <= 1e-13 per entry (different summation order and FMA)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import mprgp_amd as M
import oracle
from mprgp_amd import problems
from tests.common import brute_force_qp, kkt_report, random_spd, scene_qp

pytestmark = pytest.mark.gpu


def make_qp(ctx, A, b, node, normal, c, x0, ev=None):
    Ad = ctx.matrix_csr(A)
    q = Ad.qp()
    q.set_rows(node, normal, c)
    q.set_rhs(b)
    q.set_x(x0)
    if ev is not None:
        q.set_ev(ev)
    return q


def slack(x, node, normal, c):
    return np.einsum("ij,ij->i", normal, x.reshape(-1, 3)[node]) - c


def assert_parity(A, b, node, normal, c, x_gpu, x_ref, st_gpu, st_ref):
    rel = np.linalg.norm(x_gpu - x_ref) / np.linalg.norm(x_ref)
    assert rel <= 1e-8, rel
    f = lambda x: 0.5 * x @ (A @ x) - b @ x
    fg, fr = f(x_gpu), f(x_ref)
    assert abs(fg - fr) <= 1e-10 * abs(fr), (fg, fr)
    sg, sr = slack(x_gpu, node, normal, c), slack(x_ref, node, normal, c)
    act_g, act_r = np.abs(sg) < 1e-9, np.abs(sr) < 1e-9
    differ = act_g != act_r
    near = (np.abs(sg) < 1e-10) | (np.abs(sr) < 1e-10)
    assert not np.any(differ & ~near), np.nonzero(differ & ~near)
    assert st_gpu["code"] == st_ref["code"] == 0


def test_spmv_matches_oracle(ctx):
    A, b, node, normal, c, x0 = scene_qp(problems.config1())
    Ad = ctx.matrix_csr(A)
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.normal(size=A.shape[0])
        y = Ad.spmv(x)
        yr = oracle.spmv(A, x)
        scale = np.abs(A) @ np.abs(x)
        assert np.all(np.abs(y - yr) <= 1e-13 * scale + 1e-300)
    info = Ad.info()
    assert info["n_nodes"] == 525 and info["n_blocks"] == 525 + 2 * 2476


def test_func_value_and_diagonal_match_oracle(ctx):
    A, b, node, normal, c, x0 = scene_qp(problems.config1())
    Ad = ctx.matrix_csr(A)
    rng = np.random.default_rng(3)
    x = rng.normal(size=A.shape[0])
    f = Ad.func_value(x, b)
    fr = oracle.func_value(A, x, b)
    scale = 0.5 * np.abs(x) @ (abs(A) @ np.abs(x)) + np.abs(x) @ np.abs(b)
    assert abs(f - fr) <= 1e-14 * scale, (f, fr)
    assert np.array_equal(Ad.diagonal(), A.diagonal())


def test_bsr3_and_csr_uploads_agree(ctx):
    A, *_ = scene_qp(problems.config1())
    N = A.shape[0] // 3
    Ab = sp.bsr_matrix(A, blocksize=(3, 3))
    Ab.sort_indices()
    Mb = ctx.matrix_bsr3(N, Ab.indptr, Ab.indices, Ab.data.reshape(-1))
    Mc = ctx.matrix_csr(A)
    x = np.random.default_rng(1).normal(size=3 * N)
    np.testing.assert_array_equal(Mb.spmv(x), Mc.spmv(x))


@pytest.mark.parametrize("which", ["config1", "config1_flat"])
def test_monotone_parity_given_alpha_bar(ctx, which):
    sc = getattr(problems, which)()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    lam, _, _ = oracle.spec_rad(A, 1e-10)
    ab = 2.0 / lam
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-10, max_it=5000, alpha_bar=ab)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=1e-10, max_it=5000, alpha_bar=ab)
    xg = q.get_x()
    print(which, "oracle", sr, "gpu", sg)
    assert_parity(A, b, node, normal, c, xg, xr, sg, sr)
    assert sg["residual"] <= 1e-10 and sg["active"] == sr["active"]


@pytest.mark.parametrize("which", ["config1", "config1_flat"])
def test_monotone_same_step_sequence(ctx, which):
    # at tol 1e-6 every branch (fy <= fx, proportionality, feasibility of y)
    # is decided far from rounding, so GPU and oracle take the same steps
    sc = getattr(problems, which)()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-6, max_it=5000, alpha_bar=ab)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=1e-6, max_it=5000, alpha_bar=ab)
    print(which, "oracle", sr, "gpu", sg)
    for k in ("code", "iterations", "num_cg", "num_exp", "num_prop", "active"):
        assert sg[k] == sr[k], (k, sg, sr)
    rel = np.linalg.norm(q.get_x() - xr) / np.linalg.norm(xr)
    assert rel <= 1e-10, rel


def test_ill_conditioned_parity(ctx):
    # config 3: per-tet E log-uniform over 4 decades (kappa ~ 1e6); branch
    # decisions are rounding-sensitive, so only the solution is compared
    sc = problems.config3()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=20000, alpha_bar=ab)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=1e-8, max_it=20000, alpha_bar=ab)
    print("config3 oracle", sr, "gpu", sg)
    assert_parity(A, b, node, normal, c, q.get_x(), xr, sg, sr)


def test_power_method_matches_oracle(ctx):
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    code, xr, sr, evr = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=5000)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=1e-8, max_it=5000)
    assert sg["power_iters"] == sr["power_iters"]
    assert sg["lambda_max"] == pytest.approx(sr["lambda_max"], rel=1e-10)
    assert sg["ev_valid"] == (evr is not None)
    assert_parity(A, b, node, normal, c, q.get_x(), xr, sg, sr)
    # reference: power_ev carried into the next solve (MprgpFemSolver.cpp:533)
    if evr is not None:
        q.reset_x()
        rc2, sg2 = q.solve(tol=1e-8, max_it=5000)
        assert sg2["power_iters"] <= 2


def test_power_method_cap_matches_oracle(ctx):
    # specRad runs iterations 0..1000 when its stopping test is never met
    # (MPRGPSolver.h:58-76): the last estimate is returned and ev is left as is
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    lam, it, ev = oracle.spec_rad(A, 1e-300)
    assert it == 1001 and ev is None
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(M.params(tol=1e-8, max_it=0, power_tol=1e-300))
    assert sg["power_iters"] == 1001
    assert sg["ev_valid"] == 0
    assert sg["lambda_max"] == pytest.approx(lam, rel=1e-10)
    assert sg["alpha_bar"] == pytest.approx(2.0 / lam, rel=1e-10)


def test_mixed_power_method_uses_the_fp32_matrix(ctx):
    # the mixed form's power method multiplies by the FP32 copy of A: its
    # estimate agrees with the FP64 one to the FP32 rounding of A's values
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    lam64 = oracle.spec_rad(A, 1e-12)[0]
    lam32 = oracle.spec_rad(sp.csr_matrix(A.astype(np.float32).astype(np.float64)), 1e-12)[0]
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sm = q.solve(M.params(tol=1e-8, max_it=0, power_tol=1e-12, precision=M.MIXED))
    assert sm["lambda_max"] == pytest.approx(lam32, rel=1e-10)
    assert sm["lambda_max"] == pytest.approx(lam64, rel=1e-6)


def test_file_overload_equals_the_drop_in(ctx):
    # MPRGPDecoupledCon::solve(file_name, x, ...) (MPRGPSolver.h:576-589) reads
    # the QP and solves it exactly as the J/c overload does
    import os
    import shutil
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_tmp", "gpu_qp")
    os.makedirs(d, exist_ok=True)
    try:
        path = os.path.join(d, "config1.qp")
        M.write_qp(path, A, b, J, c, x0)
        st_f, st_m = {}, {}
        code_f, xf = M.MPRGPDecoupledCon.solve_file(path, tol=1e-8, max_it=5000, stats_out=st_f)
        xm = x0.copy()
        code_m = M.MPRGPDecoupledCon.solve(A, b, J, c, xm, tol=1e-8, max_it=5000, stats_out=st_m)
        assert code_f == code_m == 0
        assert np.array_equal(xf, xm)
        assert st_f["iterations"] == st_m["iterations"]
        code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=5000)
        assert_parity(A, b, node, normal, c, xf, xr, st_f, sr)
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("which", ["config1_flat", "config1", "config3"])
def test_classic_variant_parity(ctx, which):
    """Classic MPRGP (MPRGP::solve with the max-feasible-step search): the
    FP64 bars against the oracle on the beams, including config 3 (per-tet E
    over four decades, kappa ~1e6); the same step counts on the flat beam,
    where no decision falls within rounding of its threshold."""
    sc = getattr(problems, which)()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=5000, alpha_bar=ab,
                                             variant=oracle.CLASSIC)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=1e-8, max_it=5000, alpha_bar=ab, variant=M.CLASSIC)
    print(which, "oracle", (sr["num_cg"], sr["num_exp"], sr["num_prop"]), "gpu",
          (sg["num_cg"], sg["num_exp"], sg["num_prop"]), flush=True)
    assert code == 0 and rc == 0
    assert_parity(A, b, node, normal, c, q.get_x(), xr, sg, sr)
    if which == "config1_flat":
        assert (sg["num_cg"], sg["num_exp"], sg["num_prop"]) == (sr["num_cg"], sr["num_exp"], sr["num_prop"])


@pytest.mark.parametrize("seed", range(4))
def test_tiny_qp_matches_brute_force(ctx, seed):
    rng = np.random.default_rng(200 + seed)
    N, m = 6, 4
    A = random_spd(rng, 3 * N, 30.0) + 0.5 * np.eye(3 * N)
    b = rng.normal(size=3 * N) * 3
    node = rng.choice(N, m, replace=False).astype(np.int32)
    normal = rng.normal(size=(m, 3))
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    c = rng.normal(size=m)
    x_bf, f_bf, _ = brute_force_qp(A, b, node, normal, c)
    q = make_qp(ctx, sp.csr_matrix(A), b, node, normal, c, rng.normal(size=3 * N))
    rc, st = q.solve(tol=1e-8, max_it=10000, project_first=1)
    assert rc == 0, st
    lmin = np.linalg.eigvalsh(A)[0]
    np.testing.assert_allclose(q.get_x(), x_bf, atol=4e-8 / lmin)


def test_unconstrained_matches_cholesky(ctx):
    rng = np.random.default_rng(6)
    A = random_spd(rng, 45, 200.0) + np.eye(45)
    b = rng.normal(size=45)
    q = make_qp(ctx, sp.csr_matrix(A), b, np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0),
                np.zeros(45))
    rc, st = q.solve(tol=1e-10, max_it=1000)
    assert rc == 0
    np.testing.assert_allclose(q.get_x(), np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_reference_facade_on_test_solver_case():
    # test_case/MPRGPTest.cpp:70-98 with per-node rows (5 nodes, axis normals)
    rng = np.random.default_rng(7)
    N = 5
    Mx = rng.uniform(-1, 1, (3 * N, 3 * N)) + 3 * np.eye(3 * N)
    A = sp.csr_matrix(Mx.T @ Mx)
    node = np.arange(N, dtype=np.int32)
    normal = np.eye(3)[rng.integers(0, 3, N)]
    J = problems.rows_to_J(node, normal, N)
    B, c, x = rng.uniform(-1, 1, 3 * N), rng.uniform(-1, 1, N), rng.uniform(-1, 1, 3 * N)
    xr = x.copy()
    code_r, yr, _, _ = oracle.decoupled_solve(A, B, J, c, xr, tol=1e-8, max_it=1000)
    st = {}
    code = M.MPRGPDecoupledCon.solve(A, B, J, c, x, 1e-8, 1000, "decoupled MPRGP", None, stats_out=st)
    assert code == 0 == code_r
    assert np.all(slack(x, node, normal, c) >= -1e-9)
    np.testing.assert_allclose(x, yr, atol=1e-7)


def test_edge_cases(ctx):
    sc = problems.config1()
    A, b, node, normal, c, x0 = scene_qp(sc)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, st = q.solve(tol=1e-8, max_it=0, alpha_bar=1e-3)
    assert rc == 0 and st["iterations"] == 0            # residual_out starts at 0
    q.reset_x()
    rc, st = q.solve(tol=1e-8, max_it=3, alpha_bar=1e-3)
    assert rc == M._lib.NOT_CONVERGED and st["code"] == -1
    with pytest.raises(M.MprgpError) as e:
        q.set_rows(np.array([0, 0], np.int32), np.array([[0, 1.0, 0], [0, 1.0, 0]]), np.zeros(2))
    assert e.value.code == M._lib.EBADINPUT
    with pytest.raises(M.MprgpError):
        q.set_rows(np.array([1], np.int32), np.zeros((1, 3)), np.zeros(1))
    Jbad = sp.csr_matrix(([1.0, 1.0], ([0, 0], [0, 3])), shape=(1, A.shape[0]))  # spans 2 nodes
    with pytest.raises(M.MprgpError):
        q.set_rows_sparse(Jbad, np.zeros(1))
    # two rows with disjoint DoFs on one node (J J' diagonal, as in J = I): one row per node only
    Jbox = sp.csr_matrix(([1.0, 1.0], ([0, 1], [0, 1])), shape=(2, A.shape[0]))
    with pytest.raises(M.MprgpError) as e:
        q.set_rows_sparse(Jbox, np.zeros(2))
    assert e.value.code == M._lib.EBADINPUT and "one half-space row per node" in str(e.value)


def test_batch_equals_alone_bitwise(ctx):
    scenes = [problems.config1(seed=s) for s in (11, 12)] + [problems.config1_flat(seed=13)]
    rng = np.random.default_rng(3)
    qs, alone = [], []
    for sc in scenes:
        A, b, node, normal, c, x0 = scene_qp(sc)
        q = make_qp(ctx, A, b, node, normal, c, x0)
        rc, st = q.solve(tol=1e-9, max_it=3000, alpha_bar=2.0 / oracle.spec_rad(A, 1e-10)[0])
        alone.append((q.get_x(), st))
        qs.append((q, st["alpha_bar"]))
    for q, _ in qs:
        q.reset_x()
    per = [M.params(tol=1e-9, max_it=3000, alpha_bar=ab) for _, ab in qs]
    rc, stats = M.solve_batch(ctx, [q for q, _ in qs], per_qp=per)
    for (q, _), (xa, sa), sb in zip(qs, alone, stats):
        np.testing.assert_array_equal(q.get_x(), xa)
        for k in ("iterations", "num_cg", "num_exp", "num_prop", "active", "residual"):
            assert sb[k] == sa[k]


@pytest.mark.parametrize("prec", [M.FP64, M.MIXED])
def test_from_x0_equals_reset_then_solve(ctx, prec):
    """from_x0 = 1 restarts each QP from its stored warm start inside the
    solve (on the batch's set-up launch, or by a copy for a lone QP and the
    no-alloc solve): the same bits as mprgp_qp_reset_x and then the solve."""
    import ctypes as C
    scenes = [problems.config1(seed=s) for s in (21, 22)] + [problems.config1_flat(seed=23)]
    qs = []
    for sc in scenes:
        A, b, node, normal, c, x0 = scene_qp(sc)
        qs.append((make_qp(ctx, A, b, node, normal, c, x0), 2.0 / oracle.spec_rad(A, 1e-10)[0]))
    per = [M.params(tol=1e-8, max_it=3000, alpha_bar=ab, precision=prec, project_first=1) for _, ab in qs]
    ref = []
    for (q, _), p in zip(qs, per):
        q.reset_x()
        rc, st = q.solve(p)
        assert rc == 0
        ref.append((q.get_x(), st["iterations"]))
    # every x now holds its solution: a solve from x would stop at once
    for p in per:
        p.from_x0 = 1
    rc, stats = M.solve_batch(ctx, [q for q, _ in qs], per_qp=per)
    assert rc == 0
    for (q, _), (xr, it), s in zip(qs, ref, stats):
        assert s["iterations"] == it
        np.testing.assert_array_equal(q.get_x(), xr)
    for (q, _), (xr, it), p in zip(qs, ref, per):
        rc, st = q.solve(p)
        assert rc == 0 and st["iterations"] == it
        np.testing.assert_array_equal(q.get_x(), xr)
        st = M.Stats()
        assert M._lib.load().mprgp_qp_solve_noalloc(q.h, C.byref(p), C.byref(st)) == 0
        assert st.iterations == it
        np.testing.assert_array_equal(q.get_x(), xr)
    bad = M.Params.from_buffer_copy(per[0])
    bad.from_x0 = 2
    with pytest.raises(M.MprgpError) as e:
        qs[0][0].solve(bad)
    assert e.value.code == M._lib.EBADINPUT and "from_x0" in M.last_error()


@pytest.mark.parametrize("prec", [M.FP64, M.MIXED])
def test_index_widths_agree_bitwise(ctx, prec, monkeypatch):
    """The step kernel's SpMVs read 16-bit column indices for matrices of at
    most 65536 nodes and 32-bit ones otherwise (or with MPRGP_INDEX32=1): the
    same columns, so the same bits, in a batch and alone."""
    scenes = [problems.config1(seed=s) for s in (31, 32)] + [problems.config1_flat(seed=33)]
    qs = []
    for sc in scenes:
        A, b, node, normal, c, x0 = scene_qp(sc)
        qs.append((make_qp(ctx, A, b, node, normal, c, x0), 2.0 / oracle.spec_rad(A, 1e-10)[0]))
    per = [M.params(tol=1e-8, max_it=3000, alpha_bar=ab, precision=prec, project_first=1, from_x0=1)
           for _, ab in qs]
    out = {}
    for wide in ("0", "1"):
        monkeypatch.setenv("MPRGP_INDEX32", wide)
        rc, st = M.solve_batch(ctx, [q for q, _ in qs], per_qp=per)
        assert rc == 0
        res = [(q.get_x(), s["iterations"], s["residual"]) for (q, _), s in zip(qs, st)]
        rc, s1 = qs[2][0].solve(per[2])
        res.append((qs[2][0].get_x(), s1["iterations"], s1["residual"]))
        out[wide] = res
    for (x16, i16, r16), (x32, i32, r32) in zip(out["0"], out["1"]):
        np.testing.assert_array_equal(x16, x32)
        assert i16 == i32 and r16 == r32


def test_config2_full_size_parity(ctx):
    sc = problems.config2()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-6)[0]
    code, xr, sr, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=sc.tol, max_it=sc.max_it, alpha_bar=ab)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    rc, sg = q.solve(tol=sc.tol, max_it=sc.max_it, alpha_bar=ab)
    print("config2 oracle", sr, "gpu", sg)
    assert_parity(A, b, node, normal, c, q.get_x(), xr, sg, sr)
    rep = kkt_report(A, b, node, normal, c, q.get_x())
    assert rep["infeas"] <= 1e-12 and rep["neg_mult"] <= 1e-6


def test_copy_probe_rate_is_plausible(ctx):
    """The HBM probe behind the roofline: between a fifth of the published
    8 TB/s and the peak itself."""
    gbps = ctx.copy_probe(1 << 28, 4)
    print(f"copy probe {gbps:.0f} GB/s")
    assert 1600.0 < gbps < 8000.0


def _create_csr(ctx, n, ptr, idx, val):
    import ctypes as C
    ptr, idx, val = M._i32(ptr), M._i32(idx), M._f64(val)
    h = C.c_void_p()
    rc = M._lib.load().mprgp_matrix_create_csr(ctx.h, int(n), M._p(ptr), M._p(idx), M._p(val), C.byref(h))
    if rc == 0:
        M._lib.load().mprgp_matrix_destroy(h)
    return rc, M.last_error()


def _create_bsr3(ctx, n_nodes, rp, ci, bl):
    import ctypes as C
    rp, ci, bl = M._i32(rp), M._i32(ci), M._f64(bl)
    h = C.c_void_p()
    rc = M._lib.load().mprgp_matrix_create_bsr3(ctx.h, int(n_nodes), M._p(rp), M._p(ci), M._p(bl), C.byref(h))
    if rc == 0:
        M._lib.load().mprgp_matrix_destroy(h)
    return rc, M.last_error()


def test_malformed_matrices_are_rejected(ctx):
    """Matrix uploads check their arrays on several host threads and report
    the fault a row-by-row pass meets first: on config 2's 133,623-row matrix
    (16 ranges), a bad index late in the rows, or two faults in different
    ranges, give the message of the earlier one."""
    sc = problems.config2()
    A = scene_qp(sc)[0].tocsr()
    A.sort_indices()
    n = A.shape[0]
    ptr, idx, val = A.indptr.copy(), A.indices.copy(), A.data.copy()
    assert _create_csr(ctx, n, ptr, idx, val)[0] == 0
    late, early = ptr[120000] + 1, ptr[30000] + 2
    bad = idx.copy()
    bad[late] = n + 7
    rc, msg = _create_csr(ctx, n, ptr, bad, val)
    assert rc == M._lib.EBADINPUT and f"index {n + 7} out of range" in msg
    bad[early] = -3
    rc, msg = _create_csr(ctx, n, ptr, bad, val)
    assert rc == M._lib.EBADINPUT and "index -3 out of range" in msg
    # a decreasing ptr after an out-of-range index: the index; before it: the ptr
    p2 = ptr.copy()
    p2[90001] = p2[90000] - 1
    rc, msg = _create_csr(ctx, n, p2, bad, val)
    assert rc == M._lib.EBADINPUT and "index -3 out of range" in msg
    rc, msg = _create_csr(ctx, n, p2, idx, val)
    assert rc == M._lib.EBADINPUT and "ptr not monotone at 90000" in msg
    # block CSR: rows without their diagonal block, columns out of range,
    # non-finite values, a decreasing rowptr
    B = A.tobsr(blocksize=(3, 3))
    rp, ci, bl = B.indptr.copy(), B.indices.copy(), B.data.reshape(-1).copy()
    N = sc.n_nodes
    assert _create_bsr3(ctx, N, rp, ci, bl)[0] == 0
    c2 = ci.copy()
    row = 40000
    d = rp[row] + int(np.where(ci[rp[row]:rp[row + 1]] == row)[0][0])
    c2[d] = row + 1 if row + 1 not in ci[rp[row]:rp[row + 1]] else row - 1
    c3 = c2.copy()
    c3[rp[10000]] = N
    rc, msg = _create_bsr3(ctx, N, rp, c2, bl)
    assert rc == M._lib.EBADINPUT and f"row {row} has no diagonal block" in msg
    rc, msg = _create_bsr3(ctx, N, rp, c3, bl)
    assert rc == M._lib.EBADINPUT and f"column index {N} out of range in row 10000" in msg
    b2 = bl.copy()
    b2[9 * rp[35000] + 4] = np.nan
    b2[9 * rp[5000] + 1] = np.inf
    rc, msg = _create_bsr3(ctx, N, rp, ci, b2)
    assert rc == M._lib.EBADINPUT and f"non-finite matrix value at {9 * rp[5000] + 1}" in msg
    r2 = rp.copy()
    r2[20001] = r2[20000] - 1
    rc, msg = _create_bsr3(ctx, N, r2, ci, bl)
    assert rc == M._lib.EBADINPUT and "rowptr not monotone at row 20000" in msg
