"""Pins the oracle's plane collider (CPU only).

The reference's discrete plane collision in decoupled mode
(MprgpFemSolver.cpp:255-271 over LinearConCollider.cpp:140-161) gives each
node at most one row: that of the first plane, in the scene's order, with
n.x + p - delta <= 0 at the node's position.  These checks state that rule
by brute force in numpy and on hand-built cases (ties, infinite bands,
duplicate planes, the normal assertion), independently of the oracle's loop
over per-node plane lists."""
import math

import numpy as np
import pytest

import oracle
from mprgp_amd import problems


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def first_hit(x, planes):
    """Index of the first plane each node lies within, -1 for none (brute force)."""
    planes = np.asarray(planes, np.float64).reshape(-1, 5)
    win = np.full(x.shape[0], -1)
    for k in range(planes.shape[0] - 1, -1, -1):
        n, p, d = planes[k, :3], planes[k, 3], planes[k, 4]
        s = np.array([(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]) + p - d for v in x])
        win[s <= 0] = k
    return win


def random_planes(rng, K):
    rows = []
    for _ in range(K):
        n = unit(rng.normal(size=3))
        rows.append([*n, rng.uniform(0.3, 0.7), rng.uniform(0.0, 0.1)])  # about a third of N(0, I) each
    return np.array(rows)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rows_follow_the_first_plane_rule(seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(400, 3))
    planes = random_planes(rng, 3)
    node, normal, c = oracle.plane_collide(x, planes)
    win = first_hit(x, planes)
    assert 0 < node.size < x.shape[0]
    assert np.array_equal(np.sort(node), np.nonzero(win >= 0)[0])
    k = win[node]
    np.testing.assert_array_equal(normal, planes[k, :3])
    np.testing.assert_array_equal(c, -planes[k, 3])
    # push order: plane by plane, nodes ascending within a plane
    order = np.lexsort((node, k))
    assert np.array_equal(order, np.arange(node.size))


def test_infinite_band_takes_every_node_on_the_first_plane():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(50, 3))
    planes = [[0.0, 1.0, 0.0, 0.0, math.inf], [1.0, 0.0, 0.0, 5.0, math.inf]]
    node, normal, c = oracle.plane_collide(x, planes)
    assert np.array_equal(node, np.arange(50))
    assert np.all(normal == [0.0, 1.0, 0.0]) and np.all(c == 0.0)


def test_ties_are_inside_the_band():
    # y = 0.02 on the ground plane with delta 0.02: 0.02 + 0 - 0.02 == 0 exactly
    y_tie = 0.02
    y_out = np.nextafter(0.02, 1.0)
    x = np.array([[0.3, y_tie, -0.1], [0.3, y_out, -0.1], [-2.0, 0.0, 1.0]])
    node, normal, c = oracle.plane_collide(x, [[0.0, 1.0, 0.0, 0.0, 0.02]])
    assert list(node) == [0, 2]
    assert np.all(c == -0.0)


def test_duplicate_planes_add_no_rows():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(300, 3))
    P = random_planes(rng, 1)[0]
    one = oracle.plane_collide(x, [P])
    two = oracle.plane_collide(x, [P, P])
    for a, b in zip(one, two):
        np.testing.assert_array_equal(a, b)
    # a plane 5e-5 away (inside addConPlane's 1e-4) still gives the nodes only it reaches a row
    Q = P.copy()
    Q[3] += 5e-5
    Q[4] += 0.05
    node, normal, c = oracle.plane_collide(x, [P, Q])
    win = first_hit(x, [P, Q])
    assert np.array_equal(np.sort(node), np.nonzero(win >= 0)[0])
    assert np.any(win == 1)
    np.testing.assert_array_equal(c[win[node] == 1], -Q[3])


def test_normal_assertion():
    x = np.zeros((4, 3))
    oracle.plane_collide(x, [[0.0, 1.0 + 5e-9, 0.0, 0.0, 0.1]])
    with pytest.raises(ValueError):
        oracle.plane_collide(x, [[0.0, 1.0 + 5e-8, 0.0, 0.0, 0.1]])
    with pytest.raises(ValueError):
        oracle.plane_collide(x, [[0.0, 0.0, 0.0, 0.0, 0.1]])


def test_no_planes_and_no_nodes():
    node, normal, c = oracle.plane_collide(np.zeros((5, 3)), np.zeros((0, 5)))
    assert node.size == 0 and normal.shape == (0, 3) and c.size == 0
    node, _, _ = oracle.plane_collide(np.zeros((0, 3)), [[0.0, 1.0, 0.0, 0.0, 1.0]])
    assert node.size == 0


def test_agrees_with_the_scene_helper_on_the_wall_scene():
    sc = problems.config2()
    # move the beam 4 cm down and 1 m towards the wall so both planes take rows
    x = sc.x_n + np.array([1.0, -0.04, 0.0])
    planes = [[*unit(n), p, d] for n, p, d in sc.planes]
    node, normal, c = oracle.plane_collide(x, planes)
    hn, hnormal, hc = problems.plane_rows(x, sc.planes)
    assert node.size > 1000 and np.unique(c).size == 2
    np.testing.assert_array_equal(node, hn)
    np.testing.assert_array_equal(normal, hnormal)
    np.testing.assert_array_equal(c, hc)
