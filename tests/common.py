"""Test helpers: QPs built with the oracle's assembly, KKT checks and a
brute-force QP solver for tiny problems."""
from __future__ import annotations

import itertools
import math

import numpy as np
import scipy.sparse as sp

import oracle
from mprgp_amd import problems


def scene_qp(scene):
    """A (CSR), b, rows and warm start x0 = P(x_n) of a scene, via the oracle."""
    tets, d, vol, mass = oracle.fem_setup(scene.n_nodes, scene.tets, scene.X0)
    A, b = oracle.fem_build_system(scene.n_nodes, tets, d, vol, mass, scene.young, scene.poisson,
                                   scene.density, scene.x_n.ravel(), scene.v_n.ravel(),
                                   scene.gravity, scene.h, scene.material)
    node, normal, c = scene.rows()
    x0 = problems.project_rows(scene.x_n.ravel(), node, normal, c)
    return A, b, node, normal, c, x0


def random_spd(rng, n, cond=30.0):
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.exp(rng.uniform(0, math.log(cond), n))
    M = (Q * ev) @ Q.T
    return 0.5 * (M + M.T)


def kkt_report(A, b, node, normal, c, x):
    """Returns dict of KKT violations for min 1/2x'Ax-b'x s.t. n.x_k >= c."""
    g = A @ x - b
    X = x.reshape(-1, 3)
    G = g.reshape(-1, 3).copy()
    nrm = np.asarray(normal).reshape(-1, 3)
    s = np.einsum("ij,ij->i", nrm, X[node]) - c            # slack
    lam = np.einsum("ij,ij->i", nrm, G[node]) / np.einsum("ij,ij->i", nrm, nrm)
    # remove the normal component on rows (multiplier), leaving tangential part
    G[node] -= nrm * lam[:, None]
    free = np.ones(X.shape[0], bool)
    free[node] = False
    return {
        "infeas": float(max(0.0, -s.min())) if s.size else 0.0,
        "tangential": float(np.abs(G).max()),
        "neg_mult": float(max(0.0, -lam.min())) if lam.size else 0.0,
        "compl": float(np.abs(lam * s).max()) if s.size else 0.0,
    }


def brute_force_qp(A, b, node, normal, c):
    """Exact minimiser of a strictly convex QP with few per-node rows, by
    enumerating active sets and solving each equality-constrained KKT system."""
    A = np.asarray(A.todense()) if sp.issparse(A) else np.asarray(A)
    n = A.shape[0]
    m = len(node)
    N = np.zeros((m, n))
    for i in range(m):
        N[i, 3 * node[i]:3 * node[i] + 3] = normal[i]
    best = None
    for k in range(m + 1):
        for S in itertools.combinations(range(m), k):
            S = list(S)
            K = np.zeros((n + k, n + k))
            K[:n, :n] = A
            K[:n, n:] = -N[S].T
            K[n:, :n] = N[S]
            rhs = np.concatenate([b, c[S]])
            sol = np.linalg.solve(K, rhs)
            x, lam = sol[:n], sol[n:]
            if np.all(N @ x - c >= -1e-12) and np.all(lam >= -1e-12):
                f = 0.5 * x @ A @ x - b @ x
                if best is None or f < best[1] - 1e-14:
                    best = (x, f, S)
    return best


def truncated_counts(seq, k):
    """(num_cg, num_exp, num_prop) of a monotone solve run with max_it = k,
    from the step sequence `seq` (oracle.STEP_*) of the same solve run to
    convergence.  The truncated run takes the same first k steps; when its
    k-th step is a CG step that the full run continues with another CG step,
    the loop exits on `iteration < max_iterations` and one expansion follows
    (MPRGPSolver.h:288-290, 431-433); when the full run's next step is an
    expansion or proportioning step after a CG step, that step is taken too."""
    seq = list(seq)
    CG, EXP, PROP = 0, 1, 2
    if k >= len(seq):
        steps = seq
    elif k >= 1 and seq[k - 1] == CG:
        steps = seq[:k] + ([EXP] if seq[k] == CG else [seq[k]])
    else:
        steps = seq[:k]
    return tuple(sum(1 for s in steps if s == t) for t in (CG, EXP, PROP))


def coupled_rows(scene, rng, m):
    """Self-contact rows on a scene (problems.self_contact_rows)."""
    return problems.self_contact_rows(scene, rng, m)
