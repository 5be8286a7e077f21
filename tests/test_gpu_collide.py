"""GPU parity of the device plane collider (mprgp_qp_collide_planes).

Bar: bit-exact.  The rows it writes (node, normal, offset) equal the oracle's
plane collision at the same positions (compared by node: the oracle keeps the
reference's plane-by-plane order), and a solve after the device collider is
bitwise the solve after uploading the oracle's rows with set_rows."""
import math

import numpy as np
import pytest

import mprgp_amd as M
import oracle
from mprgp_amd import _lib, problems
from tests.common import scene_qp

pytestmark = pytest.mark.gpu


def by_node(node, normal, c):
    o = np.argsort(node, kind="stable")
    return node[o], normal[o], c[o]


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def scene_planes(sc):
    return np.array([[*unit(n), p, d] for n, p, d in sc.planes])


def check_rows(q, x, planes):
    q.set_x(np.ravel(x))
    m = q.collide_planes(planes)
    ref = by_node(*oracle.plane_collide(x, planes))
    assert m == ref[0].size
    got = q.get_rows()
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)
    return m


@pytest.fixture(scope="module")
def beam_qp(ctx):
    A, b, node, normal, c, x0 = scene_qp(problems.config1())
    q = ctx.matrix_csr(A).qp()
    q.set_rhs(b)
    return q


def test_random_planes_match_oracle(beam_qp):
    rng = np.random.default_rng(7)
    for _ in range(3):
        x = rng.normal(size=(525, 3))
        planes = np.array([[*unit(rng.normal(size=3)), rng.uniform(0.3, 0.7), rng.uniform(0.0, 0.1)]
                           for _ in range(3)])
        m = check_rows(beam_qp, x, planes)
        assert 0 < m < 525


def test_ties_bands_and_no_planes(beam_qp):
    rng = np.random.default_rng(8)
    x = rng.normal(size=(525, 3))
    x[::3, 1] = 0.02                      # n.x + p - delta == 0 exactly: inside
    x[1::3, 1] = np.nextafter(0.02, 1.0)  # one ulp outside
    ground = [0.0, 1.0, 0.0, 0.0, 0.02]
    m = check_rows(beam_qp, x, [ground])
    assert m >= 175
    assert check_rows(beam_qp, x, [[0.0, 1.0, 0.0, 0.0, math.inf]]) == 525
    # a second, duplicate plane changes nothing; then no planes clears every row
    assert check_rows(beam_qp, x, [ground, ground]) == m
    assert check_rows(beam_qp, x, np.zeros((0, 5))) == 0


def test_bad_planes_and_missing_x(ctx, beam_qp):
    x = np.zeros((525, 3))
    beam_qp.set_x(x.ravel())
    for bad in ([0.0, 1.0 + 5e-8, 0.0, 0.0, 0.1], [0.0, 0.0, 0.0, 0.0, 0.1],
                [0.0, 1.0, 0.0, math.nan, 0.1], [0.0, 1.0, 0.0, 0.0, math.nan]):
        with pytest.raises(M.MprgpError) as e:
            beam_qp.collide_planes([bad])
        assert e.value.code == _lib.EBADINPUT
    beam_qp.collide_planes([[0.0, 1.0 + 5e-9, 0.0, 0.0, 0.1]])  # within the reference's 1e-8
    A, b, *_ = scene_qp(problems.config1())
    fresh = ctx.matrix_csr(A).qp()
    with pytest.raises(M.MprgpError) as e:
        fresh.collide_planes([[0.0, 1.0, 0.0, 0.0, 0.1]])
    assert e.value.code == _lib.ESTATE


def test_wall_scene_rows_match_oracle(ctx):
    sc = problems.config2()
    A, b, *_ = scene_qp(sc)
    q = ctx.matrix_csr(A).qp()
    planes = scene_planes(sc)
    check_rows(q, sc.x_n, planes)
    x = sc.x_n + np.array([1.0, -0.04, 0.0])
    assert check_rows(q, x, planes) > 1000
    assert np.unique(q.get_rows()[2]).size == 2


@pytest.mark.parametrize("precision", [M.FP64, M.MIXED])
def test_solve_after_device_collider_is_the_host_rows_solve(ctx, precision):
    sc = problems.config2()
    A, b, *_ = scene_qp(sc)
    Ad = ctx.matrix_csr(A)
    planes = scene_planes(sc)
    # the mixed form's refinement targets the serving tolerance 1e-6
    tol = 1e-8 if precision == M.FP64 else 1e-6
    kw = dict(tol=tol, max_it=1000, project_first=1, precision=precision)
    host, dev = Ad.qp(), Ad.qp()
    for q in (host, dev):
        q.set_rhs(b)
        q.set_x(sc.x_n.ravel())
    x = sc.x_n
    for step in range(2):  # the second step collides at the first step's solution
        node, normal, c = oracle.plane_collide(x, planes)
        host.set_rows(node, normal, c)
        assert dev.collide_planes(planes) == node.size
        rh, sh = host.solve(**kw)
        rd, sd = dev.solve(**kw)
        assert rh == rd, (step, rh, rd)
        assert rh == 0, (step, sh)
        xh, xd = host.get_x(), dev.get_x()
        assert np.array_equal(xh, xd), (step, np.max(np.abs(xh - xd)))
        for k in ("iterations", "num_cg", "num_exp", "num_prop", "active"):
            assert sh[k] == sd[k], (step, k)
        x = xd.reshape(-1, 3)
