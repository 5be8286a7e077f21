"""The oracle's traced monotone solve (orc_decoupled_solve_ex) and the
closed-form objective of the device's mixed-form CG step.

The mixed form's inner solve does not evaluate f(P(y)) by a SpMV
(MPRGPSolver.h:245-249, 283-308); it forms it from the step's own sums:
f(y) = f(x) - alpha p'g + alpha^2 p'Ap/2 plus d'(A y - b) + d'A d/2 for
d = P(y) - y on the rows violated by more than 1e-9, where f(x) is the
objective at the iterate.  After an accepted step the iterate is the
unprojected y (MPRGPSolver.h:293), so f(x) continues from f(y), not from the
f(P(y)) the acceptance test compares.  The traced oracle evaluates that closed
form (and a planted wrong one continuing from f(P(y))) along the reference's
own trajectory, next to the reference's direct value; these tests pin the
correct form to the direct one on trajectories that accept infeasible trial
points, and show the planted mistake is caught."""
import numpy as np
import pytest

import oracle
from mprgp_amd import problems
from tests.common import scene_qp


def traced(scene, **kw):
    A, b, node, normal, c, x0 = scene_qp(scene)
    J = problems.rows_to_J(node, normal, scene.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    return (A, b, J, c, x0, ab), oracle.decoupled_solve_traced(A, b, J, c, x0, tol=1e-8, max_it=5000,
                                                                alpha_bar=ab, **kw)


@pytest.mark.parametrize("name", ["config1", "config1_flat", "config3"])
def test_tracing_changes_nothing(name):
    (A, b, J, c, x0, ab), (code, x, st, ev) = traced(getattr(problems, name)())
    code0, x0r, st0, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=1e-8, max_it=5000, alpha_bar=ab,
                                                project_first=True)
    assert code == code0 and st == st0 and np.array_equal(x, x0r)
    seq = oracle.step_sequence(ev)
    assert len(seq) == st["iterations"]
    assert [int(np.sum(seq == k)) for k in (oracle.STEP_CG, oracle.STEP_EXP, oracle.STEP_PROP)] == \
        [st["num_cg"], st["num_exp"], st["num_prop"]]


@pytest.mark.parametrize("name", ["config1", "config1_flat", "config3"])
def test_closed_form_objective_matches_the_direct_one(name):
    _, (code, x, st, ev) = traced(getattr(problems, name)())
    k = np.where(ev["kind"] == oracle.EV_CG_TEST)[0]
    assert np.all(ev["kind"][k + 1] == oracle.EV_FY_CLOSED)
    fy, fix, bad = ev["lhs"][k], ev["lhs"][k + 1], ev["rhs"][k + 1]
    # accepted trial points that violate a plane: the case the closed form's
    # continuation has to get right
    nxt = np.minimum(k + 2, len(ev) - 1)
    accepted_infeasible = int(np.sum((ev["outcome"][k] == 1) & (ev["kind"][nxt] == oracle.EV_Y_INFEASIBLE)))
    assert accepted_infeasible >= 5, accepted_infeasible
    rel_fix = np.abs(fix - fy) / np.abs(fy)
    rel_bad = np.abs(bad - fy) / np.abs(fy)
    assert rel_fix.max() <= 1e-13, rel_fix.max()
    # the planted continuation from f(P(y)) drifts by the projections' shares
    assert rel_bad.max() >= 1e-8, rel_bad.max()


def test_closed_form_decisions_converge_to_the_same_solution():
    """Decisions taken on the closed form (the mixed form's rule) reach the
    reference's solution; the closed form resolves near-ties of fy <= fx
    that the direct difference of two large sums decides by rounding, so it
    takes fewer expansion steps."""
    sc = problems.config1()
    (A, b, J, c, x0, ab), (code_d, x_d, st_d, _) = traced(sc)
    code_c, x_c, st_c, ev_c = oracle.decoupled_solve_traced(A, b, J, c, x0, tol=1e-8, max_it=5000,
                                                            alpha_bar=ab, fmode=oracle.F_CLOSED)
    assert code_d == 0 and code_c == 0
    assert np.linalg.norm(x_c - x_d) <= 1e-8 * np.linalg.norm(x_d)
    assert st_c["num_exp"] <= st_d["num_exp"]
