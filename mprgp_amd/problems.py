"""Synthetic contact-QP scenes for the solver's workloads.

The reference reads ABQ meshes and JSON scenes (Simulator.cpp:29-359); the
workloads here are generated instead, from a seed:

* ``box_mesh``: hex grid split into 5 tets per hex with the split alternating
  by cell parity so face diagonals conform (18 edges per hex, no body
  diagonal); node index ``(i*(ny+1)+j)*(nz+1)+k`` keeps the long x axis
  slowest, which keeps neighbour columns close in memory.
* ``place``: rotation about z, then a lift so that min y equals ``lift``;
  applied to rest and current positions alike, as the reference's
  ``applyTrans(R, i, true, true)`` does (FEMMesh.cpp:245-254).
* ``plane_rows``: LinearConCollider::handle(plane) semantics
  (LinearConCollider.cpp:140-161): a node gets a row for plane (n, p) when
  ``n.x + p - delta <= 0``; in decoupled mode only its first such plane makes
  a row (later planes are recorded but add none); the row is ``n . x_k >= -p``
  (GeomConCache::getConstraints, :58-72).

Random draws use numpy's PCG64 (``numpy.random.default_rng(seed)``).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_EVEN = np.array([[1, 0, 3, 5], [2, 0, 3, 6], [4, 0, 5, 6], [7, 3, 5, 6], [0, 3, 5, 6]])
_ODD = np.array([[0, 1, 2, 4], [3, 1, 2, 7], [5, 1, 4, 7], [6, 2, 4, 7], [1, 2, 4, 7]])


def box_mesh(nx: int, ny: int, nz: int, a: float):
    """Returns (X[N,3], tets[T,4]) of an nx*ny*nz hex grid with edge a."""
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([I, J, K], -1).reshape(-1, 3).astype(np.float64) * a
    ci, cj, ck = [v.ravel() for v in np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                                                  indexing="ij")]

    def nid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    corners = np.stack([nid(ci + (b & 1), cj + ((b >> 1) & 1), ck + ((b >> 2) & 1))
                        for b in range(8)], -1)  # [cells, 8]
    even = ((ci + cj + ck) % 2) == 0
    tets = np.where(even[:, None, None], corners[:, _EVEN], corners[:, _ODD]).reshape(-1, 4)
    return X, orient_positive(X, tets)


def signed_volumes(X, tets):
    a, b, c, d = (X[tets[:, i]] for i in range(4))
    return np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) / 6.0


def orient_positive(X, tets):
    """makePositive (FEMMesh.cpp:158-165): swap v2, v3 of negative tets."""
    tets = tets.copy()
    neg = signed_volumes(X, tets) < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return tets


def rot_z(deg):
    t = math.radians(deg)
    return np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1.0]])


def random_rotation(rng):
    """Uniform random rotation from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def place(X, angle_z_deg=0.0, lift=0.0, shift_x=0.0):
    Y = X @ rot_z(angle_z_deg).T
    Y[:, 1] += lift - Y[:, 1].min()
    Y[:, 0] += shift_x
    return Y


def plane_rows(x, planes):
    """planes: list of (normal[3], p, delta).  Returns (node, normal[m,3], c[m])."""
    N = x.shape[0]
    taken = np.zeros(N, dtype=bool)
    nodes, normals, cs = [], [], []
    for n, p, delta in planes:
        n = np.asarray(n, dtype=np.float64)
        n = n / np.linalg.norm(n)
        hit = (x @ n + p - delta) <= 0
        new = np.nonzero(hit & ~taken)[0]
        taken[new] = True
        nodes.append(new)
        normals.append(np.repeat(n[None, :], new.size, 0))
        cs.append(np.full(new.size, -float(p)))
    if not nodes:
        return np.zeros(0, np.int32), np.zeros((0, 3)), np.zeros(0)
    return (np.concatenate(nodes).astype(np.int32), np.concatenate(normals, 0),
            np.concatenate(cs))


def plane_array(planes):
    """(K, 5) {nx, ny, nz, p, delta} of a scene's (normal, p, delta) planes,
    normals scaled to unit length as ``plane_rows`` uses them."""
    out = np.zeros((len(planes), 5))
    for k, (n, p, delta) in enumerate(planes):
        n = np.asarray(n, dtype=np.float64)
        out[k, :3] = n / np.linalg.norm(n)
        out[k, 3] = p
        out[k, 4] = delta
    return out


def rows_to_J(node, normal, n_nodes):
    """Sparse J (m x 3N) of per-node rows, as LinearConCollider::getConstraints builds it."""
    import scipy.sparse as sp
    m = node.size
    rows = np.repeat(np.arange(m), 3)
    cols = (3 * node[:, None] + np.arange(3)[None, :]).ravel()
    return sp.csr_matrix((normal.ravel(), (rows, cols)), shape=(m, 3 * n_nodes))


def project_rows(x, node, normal, c):
    """DecoupledConProjector::project in per-node form (numpy, host-side)."""
    y = x.reshape(-1, 3).copy()
    nrm = np.asarray(normal).reshape(-1, 3)
    jj = np.einsum("ij,ij->i", nrm, nrm)
    lam = np.maximum((c - np.einsum("ij,ij->i", nrm, y[node])) / jj, 0.0)
    y[node] += nrm * lam[:, None]
    return y.ravel()


def self_contact_rows(scene, rng, m, gap=0.0):
    """m coupled rows of 12 nonzeros in the reference's self-contact form
    (SurfaceSelfConCache, LinearConCollider.cpp:74-126 and :186-217): a vertex
    i against the triangle (j, k, l) of one tet face, unit normal n of that
    face oriented towards i, barycentric weights (a, b, c):
    n.(x_i - a x_j - b x_k - c x_l) >= rhs, with rhs the row's value at the
    scene's current positions minus `gap` (so the warm start is feasible and
    the dynamics press on about half of the rows).  Distinct vertices, and no
    vertex also in another row's triangle.  Returns (J as an m x 3N CSR, rhs)."""
    import scipy.sparse as sp
    x = scene.x_n
    N = scene.n_nodes
    used = np.zeros(N, bool)
    rows, cols, vals, rhs = [], [], [], []
    faces = [(0, 1, 2, 3), (0, 1, 3, 2), (0, 2, 3, 1), (1, 2, 3, 0)]
    for t in rng.permutation(len(scene.tets)):
        if len(rhs) == m:
            break
        j, k, l, i = (int(scene.tets[t][q]) for q in faces[int(rng.integers(0, 4))])
        if used[[i, j, k, l]].any():
            continue
        nrm = np.cross(x[k] - x[j], x[l] - x[j])
        nrm /= np.linalg.norm(nrm)
        if nrm @ (x[i] - x[j]) < 0.0:
            nrm = -nrm
        w = rng.dirichlet([3.0, 3.0, 3.0])
        r = len(rhs)
        for node, coef in ((i, 1.0), (j, -w[0]), (k, -w[1]), (l, -w[2])):
            for d in range(3):
                rows.append(r)
                cols.append(3 * node + d)
                vals.append(coef * nrm[d])
        rhs.append(float(nrm @ (x[i] - w[0] * x[j] - w[1] * x[k] - w[2] * x[l])) - gap)
        used[[i, j, k, l]] = True
    J = sp.csr_matrix((vals, (rows, cols)), shape=(len(rhs), 3 * N))
    J.sort_indices()
    return J, np.array(rhs)


@dataclasses.dataclass
class Scene:
    """One body on planes: everything buildLinearSystem + the collider need."""
    name: str
    X0: np.ndarray          # rest positions (N,3), placed
    tets: np.ndarray        # (T,4) positively oriented
    young: np.ndarray       # (T,)
    poisson: np.ndarray     # (T,)
    density: float
    x_n: np.ndarray         # current positions (N,3)
    v_n: np.ndarray         # velocities (N,3)
    gravity: np.ndarray
    h: float
    material: int           # 1 LINEAR, 2 COROTATIONAL
    planes: list
    tol: float
    max_it: int

    @property
    def n_nodes(self):
        return self.X0.shape[0]

    def rows(self):
        return plane_rows(self.x_n, self.planes)


def beam_scene(name, nx, ny, nz, a=0.05, young=1e5, poisson=0.45, density=1000.0, h=0.01,
               angle=5.0, lift=0.01, v0=-1.0, v_sigma=0.05, planes=None, tol=1e-8, max_it=1000,
               material=1, seed=0, young_range=None, rng=None):
    rng = np.random.default_rng(seed) if rng is None else rng
    X, tets = box_mesh(nx, ny, nz, a)
    X0 = place(X, angle, lift)
    tets = orient_positive(X0, tets)
    T = tets.shape[0]
    if young_range is not None:
        lo, hi = young_range
        E = np.exp(rng.uniform(math.log(lo), math.log(hi), T))
    else:
        E = np.full(T, float(young))
    nu = np.full(T, float(poisson)) if np.isscalar(poisson) else np.asarray(poisson, np.float64)
    v = np.zeros_like(X0)
    v[:, 1] = v0
    v += rng.normal(0.0, v_sigma, v.shape)
    if planes is None:
        planes = [((0.0, 1.0, 0.0), 0.0, math.inf)]
    return Scene(name, X0, tets, E, nu, density, X0.copy(), v, np.array([0.0, -9.8, 0.0]), h,
                 material, planes, tol, max_it)


def config1(seed=1):
    """Linear tet beam 20x4x4 cells (525 nodes, 1600 tets), rotated 5 degrees
    about z with its lowest node 1 cm above y = 0, y >= 0 on every node."""
    return beam_scene("beam20x4x4", 20, 4, 4, seed=seed, tol=1e-8, max_it=1000)


def config1_flat(seed=1):
    """Config 1 dropped flat (no tilt): the whole bottom face (20% of the
    nodes) ends on the ground."""
    return beam_scene("beam20x4x4-flat", 20, 4, 4, seed=seed, angle=0.0, tol=1e-8, max_it=1000)


def config2(seed=2):
    """100x20x20 cells (44,541 nodes, 200,000 tets) on the ground y >= 0 and a
    wall through x = 4.8 tilted 10 degrees, rows inside a 2 cm band."""
    t = math.radians(10.0)
    n_wall = (-math.cos(t), math.sin(t), 0.0)
    p_wall = 4.8 * math.cos(t)
    planes = [((0.0, 1.0, 0.0), 0.0, 0.02), (n_wall, p_wall, 0.02)]
    return beam_scene("beam100x20x20", 100, 20, 20, seed=seed, angle=0.0, lift=0.005,
                      planes=planes, tol=1e-8, max_it=500)


def config3(seed=3):
    """Config-1 mesh with per-tet Young's modulus log-uniform in [1e3, 1e7] Pa."""
    return beam_scene("beam20x4x4-hetero", 20, 4, 4, seed=seed, young_range=(1e3, 1e7),
                      tol=1e-8, max_it=5000)


def dims_for_dof(dof, aspect=5.0):
    """Grid (nx, k, k) whose node count (nx+1)(k+1)^2 is close to dof/3."""
    nodes = dof / 3.0
    k = max(2, int(round((nodes / aspect) ** (1.0 / 3.0))) - 1)
    nx = max(2, int(round(nodes / (k + 1) ** 2)) - 1)
    return nx, k, k


def random_planes(rng, x, n_planes, frac_total, delta=0.01):
    """Planes with normals within 30 deg of +y whose offsets cut `frac_total`
    of the nodes in total (split evenly over the planes)."""
    planes = []
    for _ in range(n_planes):
        tilt = math.radians(rng.uniform(0.0, 30.0))
        az = rng.uniform(0.0, 2 * math.pi)
        n = np.array([math.sin(tilt) * math.cos(az), math.cos(tilt), math.sin(tilt) * math.sin(az)])
        s = x @ n
        p = -float(np.quantile(s, frac_total / n_planes))
        planes.append((tuple(n), p, delta))
    return planes


def batch_scenes(count=256, seed=4, dof_range=(5e3, 6e4), tol=1e-6, max_it=2000):
    """Independent beams, DoF log-uniform in dof_range, E log-uniform [1e4, 1e6],
    nu uniform [0.3, 0.48], 1-4 planes placing 2-15% of the nodes on them."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dof = math.exp(rng.uniform(math.log(dof_range[0]), math.log(dof_range[1])))
        nx, ny, nz = dims_for_dof(dof)
        E = math.exp(rng.uniform(math.log(1e4), math.log(1e6)))
        nu = rng.uniform(0.3, 0.48)
        sc = beam_scene(f"batch{i}", nx, ny, nz, young=E, poisson=nu, angle=rng.uniform(0, 10),
                        tol=tol, max_it=max_it, rng=rng)
        k = int(rng.integers(1, 5))
        frac = rng.uniform(0.02, 0.15)
        sc.planes = random_planes(rng, sc.x_n, k, frac)
        out.append(sc)
    return out
