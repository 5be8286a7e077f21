"""Step sequences of the FP64 device solve against the oracle's.

The device solve follows the reference's control flow (MPRGPSolver.h:
272-316, 412-454) decision by decision, but its sums are not the oracle's:
dot products and the objective 1/2 y'Ay - y'b reduce in a fixed tree
(per-lane 3-vectors, DPP within the wavefront, then waves and tiles) where the
oracle sums left to right, and the SpMV rounds with fused multiply-adds.
Where a decision is within rounding of its threshold, the two may branch
differently; from there on the sequences differ while both still converge
to the same minimiser.

Each case runs the device solve with max_it = k for every k (the truncated
runs take the same first steps as the full run; tests/common.py's
truncated_counts gives the oracle's counts for the same k from its traced
full run) and finds the first k where the step counts differ.  At that point
the oracle's own trace must show a decision within rounding of its
threshold: the monotone test fy <= fx (MPRGPSolver.h:288), the proportioning
test (:174-181) or the feasibility of y (:431), at the iterations k-1..k+1,
with a relative margin below 1e-12 on the well-conditioned beams.  On the
ill-conditioned beam (per-tet E over four decades, kappa ~ 1e6) the iterates
themselves drift apart by the accumulated rounding of many CG steps (printed:
the distance of the two iterates just before the flip), and the first flip
comes at a CG step that changes the objective by ~1e-10 of its value; the
bar there is 1e-9.  Cases that never diverge pass with the same sequence
throughout."""
import numpy as np
import pytest

import oracle
from mprgp_amd import problems
from tests.common import scene_qp, truncated_counts
from tests.test_gpu_solver import make_qp

pytestmark = pytest.mark.gpu

TIE = {"config3": 1e-9}  # default 1e-12


def margins(events, it_lo, it_hi):
    """(kind, relative margin) of the oracle's decisions at iterations it_lo..it_hi."""
    out = []
    for e in events:
        if not (it_lo <= e["iteration"] <= it_hi):
            continue
        k, lhs, rhs = int(e["kind"]), float(e["lhs"]), float(e["rhs"])
        if k == oracle.EV_CG_TEST:
            out.append(("fy<=fx", abs(lhs - rhs) / max(abs(rhs), 1e-300)))
        elif k == oracle.EV_PROP:
            out.append(("proportional", abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)))
        elif k == oracle.EV_FEASIBLE:
            out.append(("feasible(y)", abs(lhs - rhs)))
    return out


def first_divergence(q, seq, tol, ab, k_max):
    for k in range(1, k_max + 1):
        q.reset_x()
        rc, st = q.solve(tol=tol, max_it=k, alpha_bar=ab, project_first=1)
        got = (st["num_cg"], st["num_exp"], st["num_prop"])
        if got != truncated_counts(seq, k):
            return k, got
    return None, None


@pytest.mark.parametrize("which,tol", [("config1", 1e-6), ("config1_flat", 1e-6), ("config1", 1e-8),
                                       ("config3", 1e-8), ("config2", 1e-6)])
def test_first_divergence_is_a_rounding_tie(ctx, which, tol):
    sc = getattr(problems, which)()
    A, b, node, normal, c, x0 = scene_qp(sc)
    J = problems.rows_to_J(node, normal, sc.n_nodes)
    ab = 2.0 / oracle.spec_rad(A, 1e-10)[0]
    max_it = 500 if which == "config2" else 5000
    code, xr, sr, events = oracle.decoupled_solve_traced(A, b, J, c, x0, tol=tol, max_it=max_it, alpha_bar=ab)
    seq = oracle.step_sequence(events)
    q = make_qp(ctx, A, b, node, normal, c, x0)
    k, got = first_divergence(q, seq, tol, ab, len(seq) + 1)
    q.reset_x()
    rc, sg = q.solve(tol=tol, max_it=max_it, alpha_bar=ab, project_first=1)
    print(f"{which} tol {tol}: oracle {len(seq)} steps {sr['num_cg']}/{sr['num_exp']}/{sr['num_prop']}, "
          f"device {sg['iterations']} steps {sg['num_cg']}/{sg['num_exp']}/{sg['num_prop']}")
    assert rc == 0 and code == 0
    rel = np.linalg.norm(q.get_x() - xr) / np.linalg.norm(xr)
    assert rel <= (1e-8 if tol <= 1e-8 else 1e-6), rel
    if k is None:
        print("  same step sequence throughout")
        assert (sg["num_cg"], sg["num_exp"], sg["num_prop"]) == (sr["num_cg"], sr["num_exp"], sr["num_prop"])
        return
    m = margins(events, k - 1, k + 1)
    tight = min(m, key=lambda t: t[1])
    q.reset_x()
    q.solve(tol=tol, max_it=k - 1, alpha_bar=ab, project_first=1)
    _, xo, _, _ = oracle.decoupled_solve(A, b, J, c, x0, tol=tol, max_it=k - 1, alpha_bar=ab, project_first=True)
    drift = np.linalg.norm(q.get_x() - xo) / np.linalg.norm(xo)
    print(f"  first divergence at max_it = {k}: device {got}, oracle {truncated_counts(seq, k)}; "
          f"tightest oracle decision there: {tight[0]} at relative margin {tight[1]:.2e}; "
          f"iterates {drift:.1e} apart just before")
    assert tight[1] <= TIE.get(which, 1e-12), m
