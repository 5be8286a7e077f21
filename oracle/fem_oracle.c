/*
 * fem_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Serial FP64 restatement of the reference's FEM system assembly for the
 * LINEAR and COROTATIONAL materials (the two the contact QP path uses).
 *
 * Function -> reference (file:line) -> how it is pinned (tests/test_oracle_fem.py):
 *   orc_fem_setup        FEMMesh.cpp:158-184, 488-510 (makePositive, buildD, assemble)
 *                        -> volumes equal |det|/6, sum of node masses equals total volume,
 *                           d * Dm = I
 *   orc_polar_rotation   FEMMesh.cpp:123-157 + FEMRotationUtil.h:959-1010 (adjustF3D)
 *                        -> R orthogonal with det +1, R^T F symmetric positive
 *                           semi-definite, R = Q for F = Q S with known Q, S
 *   orc_fem_element      FEMEnergy.cpp:263-304 (MaterialEnergy::eval), 377-394 (setParam),
 *                        657-800 (evalFLinear), 801-1057 (evalFCorot); FEMUtils.h:30-41 (calcGComp3D)
 *                        -> LINEAR K_e equals the textbook V B^T D B (Voigt form written
 *                           independently in the test); rigid translations and infinitesimal
 *                           rotations in the null space; finite differences of the energy give
 *                           the force and of the force give the stiffness; COROT with R = I
 *                           equals LINEAR; rotating the element by Q gives K' = (I x Q) K (I x Q)^T
 *   orc_fem_build_system FEMSystem.cpp:167-192, 558-569; FEMEnergy.cpp:98-115;
 *                        MprgpFemSolver.cpp:295-356
 *                        -> for LINEAR, b equals the closed form M(x_n + h v_n) + h^2(K X0 + M g);
 *                           A symmetric; A - M = h^2 sum K_e; sum of masses = density * volume
 *   orc_fem_build_system_at  the same at a Newton iterate x1 (MprgpFemSolver.cpp:315-356)
 *                        (tests/test_oracle_step.py)
 *                        -> for LINEAR, b does not depend on x1 (the closed form above); for
 *                           COROT, b - A x1 - M(x_n + h v_n - x1) - h^2 M g equals h^2 times
 *                           minus the finite-difference gradient of the total energy at x1
 *   orc_fem_build_system_damped  the same with Rayleigh damping (FEMSystem.cpp:122-126, 167-192)
 *                        -> A = M (1 + h alpha_k) + (h^2 + h alpha_m)/h^2 (A_0 - M) against the
 *                           undamped A_0; the damping terms of b against the undamped b; a
 *                           damped free vibration meets implicit Euler's discrete energy identity
 *                           E_{n+1} - E_n = -|dv|_M^2/2 - |du|_K^2/2 - h v' (alpha_k M + alpha_m K) v
 *   orc_implicit_euler_velocity  MprgpFemSolver.cpp:295-313, 377-390 with FEMSystem.h:140-142
 *                        -> equals (x - x_n)/h to rounding; a rigid free fall of one step from
 *                           the rest shape gives x = X0 + h v + h^2 g and v' = v + h g
 */
#include "mprgp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

static double det3(const double *M) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
         M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* inverse of a row-major 3x3 by the adjugate */
static void inv3(const double *M, double *I) {
  double c00 = M[4] * M[8] - M[5] * M[7];
  double c01 = M[5] * M[6] - M[3] * M[8];
  double c02 = M[3] * M[7] - M[4] * M[6];
  double det = M[0] * c00 + M[1] * c01 + M[2] * c02;
  double id = 1.0 / det;
  I[0] = c00 * id;
  I[1] = (M[2] * M[7] - M[1] * M[8]) * id;
  I[2] = (M[1] * M[5] - M[2] * M[4]) * id;
  I[3] = c01 * id;
  I[4] = (M[0] * M[8] - M[2] * M[6]) * id;
  I[5] = (M[2] * M[3] - M[0] * M[5]) * id;
  I[6] = c02 * id;
  I[7] = (M[1] * M[6] - M[0] * M[7]) * id;
  I[8] = (M[0] * M[4] - M[1] * M[3]) * id;
}

/* signed volume (b-a)x(c-a).(d-a)/6, TetrahedronTpl::volume */
static double tet_volume(const double *a, const double *b, const double *c, const double *d) {
  double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  double x[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  return (x[0] * w[0] + x[1] * w[1] + x[2] * w[2]) / 6.0;
}

void orc_fem_setup(int n_nodes, int n_tets, int *tets, const double *X0,
                   double *d, double *vol, double *node_mass) {
  for (int i = 0; i < n_nodes; ++i) node_mass[i] = 0.0;
  for (int t = 0; t < n_tets; ++t) {
    int *v = tets + 4 * t;
    /* makePositive: swap v2, v3 when the signed volume is negative */
    if (tet_volume(X0 + 3 * v[0], X0 + 3 * v[1], X0 + 3 * v[2], X0 + 3 * v[3]) < 0.0) {
      int s = v[2]; v[2] = v[3]; v[3] = s;
    }
    vol[t] = tet_volume(X0 + 3 * v[0], X0 + 3 * v[1], X0 + 3 * v[2], X0 + 3 * v[3]);
    /* buildD: columns are X1-X0, X2-X0, X3-X0 of the rest shape */
    double Dm[9];
    for (int r = 0; r < 3; ++r)
      for (int i = 0; i < 3; ++i) Dm[r * 3 + i] = X0[3 * v[i + 1] + r] - X0[3 * v[0] + r];
    inv3(Dm, d + 9 * t);
  }
  for (int t = 0; t < n_tets; ++t)
    for (int j = 0; j < 4; ++j) node_mass[tets[4 * t + j]] += vol[t] / 4.0;
}

/* cyclic Jacobi eigen-decomposition of a symmetric 3x3 (row-major A, destroyed);
 * eigenvalues ascending in w, eigenvectors as columns of V (row-major). */
static void jacobi3(double *A, double *w, double *V) {
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    double diag = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
    if (off <= 1e-40 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double apq = A[p * 3 + q];
        if (apq == 0.0) continue;
        double app = A[p * 3 + p], aqq = A[q * 3 + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) { /* A = A J */
          double akp = A[k * 3 + p], akq = A[k * 3 + q];
          A[k * 3 + p] = c * akp - s * akq;
          A[k * 3 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) { /* A = J^T A */
          double apk = A[p * 3 + k], aqk = A[q * 3 + k];
          A[p * 3 + k] = c * apk - s * aqk;
          A[q * 3 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) { /* V = V J */
          double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
          V[k * 3 + p] = c * vkp - s * vkq;
          V[k * 3 + q] = s * vkp + c * vkq;
        }
      }
  }
  w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
  for (int i = 0; i < 2; ++i) /* sort ascending, permuting columns */
    for (int j = 0; j < 2 - i; ++j)
      if (w[j] > w[j + 1]) {
        double t = w[j]; w[j] = w[j + 1]; w[j + 1] = t;
        for (int k = 0; k < 3; ++k) {
          double tv = V[k * 3 + j]; V[k * 3 + j] = V[k * 3 + j + 1]; V[k * 3 + j + 1] = tv;
        }
      }
}

void orc_polar_rotation(const double *F, double *R) {
  double C[9], w[3], V[9], U[9], S[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += F[k * 3 + i] * F[k * 3 + j];
      C[i * 3 + j] = s;
    }
  jacobi3(C, w, V);
  for (int i = 0; i < 3; ++i) S[i] = sqrt(fabs(w[i]));
  if (det3(V) < 0.0)
    for (int k = 0; k < 3; ++k) V[k * 3 + 0] = -V[k * 3 + 0];
  const double F_EPS = (double)1e-6f;
  int good[3], nr = 0;
  for (int v = 0; v < 3; ++v)
    if (S[v] > F_EPS) {
      for (int r = 0; r < 3; ++r) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += F[r * 3 + k] * V[k * 3 + v];
        U[r * 3 + v] = s / S[v];
      }
      good[nr++] = v;
    }
  if (nr == 0) {
    for (int i = 0; i < 9; ++i) U[i] = (i % 4 == 0) ? 1.0 : 0.0;
  } else {
    for (int v = 0; v < 3; ++v) {
      if (S[v] > F_EPS) continue;
      double a[3] = {U[0 * 3 + good[0]], U[1 * 3 + good[0]], U[2 * 3 + good[0]]};
      double b[3];
      if (nr == 1) {
        /* the reference crosses with a random vector; a fixed axis least
           aligned with a keeps this deterministic */
        int ax = 0;
        if (fabs(a[1]) < fabs(a[ax])) ax = 1;
        if (fabs(a[2]) < fabs(a[ax])) ax = 2;
        b[0] = b[1] = b[2] = 0.0; b[ax] = 1.0;
      } else {
        b[0] = U[0 * 3 + good[1]]; b[1] = U[1 * 3 + good[1]]; b[2] = U[2 * 3 + good[1]];
      }
      double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      double nd = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
      if (nr == 1) { cx[0] /= nd; cx[1] /= nd; cx[2] /= nd; }
      for (int r = 0; r < 3; ++r) U[r * 3 + v] = cx[r];
      good[nr++] = v;
    }
  }
  if (det3(U) < 0.0) {
    int id = 0;
    if (S[1] < S[id]) id = 1;
    if (S[2] < S[id]) id = 2;
    for (int k = 0; k < 3; ++k) U[k * 3 + id] = -U[k * 3 + id];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += U[i * 3 + k] * V[j * 3 + k];
      R[i * 3 + j] = s;
    }
}

/* calcGComp3D (FEMUtils.h:30-41): 9x12, vec(F) index c*3+r, x_e node-major */
static void gcomp(const double *d, double *G) {
  memset(G, 0, sizeof(double) * 108);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < 3; ++i) {
        G[(c * 3 + r) * 12 + (i + 1) * 3 + r] += d[i * 3 + c];
        G[(c * 3 + r) * 12 + r] -= d[i * 3 + c];
      }
}

double orc_fem_element(int type, const double *xe, const double *d, double V,
                       double lambda, double mu, double *fB, double *HB) {
  double F[9], R[9], grad[9], H[81], G[108];
  /* FEMCell::buildF: F = DX * d with DX columns x_{i+1} - x_0 */
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += (xe[(i + 1) * 3 + r] - xe[r]) * d[i * 3 + c];
      F[r * 3 + c] = s;
    }
  if (type == ORC_COROTATIONAL) {
    orc_polar_rotation(F, R);
  } else {
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  /* S = R^T F (LINEAR: S = F), eps = sym(S) - I */
  double S[9], eps[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (type == ORC_COROTATIONAL) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += R[k * 3 + i] * F[k * 3 + j];
        S[i * 3 + j] = s;
      } else {
        S[i * 3 + j] = F[i * 3 + j];
      }
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) eps[i * 3 + j] = 0.5 * (S[i * 3 + j] + S[j * 3 + i]) - (i == j ? 1.0 : 0.0);
  double tr = eps[0] + eps[4] + eps[8];
  double ee = 0.0;
  for (int i = 0; i < 9; ++i) ee += eps[i] * eps[i];
  double E = V * (0.5 * lambda * tr * tr + mu * ee);
  /* grad_F = V R (lambda tr(eps) I + 2 mu eps), stored col-major (c*3+r) */
  double P[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P[i * 3 + j] = lambda * tr * (i == j ? 1.0 : 0.0) + 2.0 * mu * eps[i * 3 + j];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += R[r * 3 + k] * P[k * 3 + c];
      grad[c * 3 + r] = V * s;
    }
  /* Hessian in F with R held fixed:
     V [ lambda R_rc R_r'c' + mu ( (R R^T)_rr' delta_cc' + R_rc' R_r'c ) ] */
  double RRt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += R[i * 3 + k] * R[j * 3 + k];
      RRt[i * 3 + j] = s;
    }
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      for (int c2 = 0; c2 < 3; ++c2)
        for (int r2 = 0; r2 < 3; ++r2) {
          double h = lambda * R[r * 3 + c] * R[r2 * 3 + c2] +
                     mu * ((c == c2 ? RRt[r * 3 + r2] : 0.0) + R[r * 3 + c2] * R[r2 * 3 + c]);
          H[(c * 3 + r) * 9 + c2 * 3 + r2] = V * h;
        }
  gcomp(d, G);
  /* fB = -grad^T G ; HB = G^T H G */
  for (int j = 0; j < 12; ++j) {
    double s = 0.0;
    for (int a = 0; a < 9; ++a) s += grad[a] * G[a * 12 + j];
    fB[j] = -s;
  }
  double HG[108];
  for (int a = 0; a < 9; ++a)
    for (int j = 0; j < 12; ++j) {
      double s = 0.0;
      for (int b = 0; b < 9; ++b) s += H[a * 9 + b] * G[b * 12 + j];
      HG[a * 12 + j] = s;
    }
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0.0;
      for (int a = 0; a < 9; ++a) s += G[a * 12 + i] * HG[a * 12 + j];
      HB[i * 12 + j] = s;
    }
  return E;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int orc_fem_build_system_damped(int n_nodes, int n_tets, const int *tets,
                                const double *d, const double *vol, const double *node_mass,
                                const double *young, const double *poisson, double density,
                                const double *x_n, const double *v_n, const double *x1,
                                const double *gravity, double h, int type, double alpha_k,
                                double alpha_m, int **ptr_out, int **idx_out, double **val_out,
                                double *b) {
  /* node adjacency (including self), sorted */
  int *deg = (int *)xcalloc((size_t)n_nodes + 1, sizeof(int));
  for (int t = 0; t < n_tets; ++t)
    for (int i = 0; i < 4; ++i) deg[tets[4 * t + i]] += 4;
  int *aptr = (int *)xcalloc((size_t)n_nodes + 1, sizeof(int));
  for (int i = 0; i < n_nodes; ++i) aptr[i + 1] = aptr[i] + deg[i];
  int *adj = (int *)xcalloc((size_t)aptr[n_nodes], sizeof(int));
  int *fill = (int *)xcalloc((size_t)n_nodes, sizeof(int));
  for (int t = 0; t < n_tets; ++t)
    for (int i = 0; i < 4; ++i) {
      int a = tets[4 * t + i];
      for (int j = 0; j < 4; ++j) adj[aptr[a] + fill[a]++] = tets[4 * t + j];
    }
  int *nptr = (int *)xcalloc((size_t)n_nodes + 1, sizeof(int));
  int *nadj = (int *)xcalloc((size_t)aptr[n_nodes] + 1, sizeof(int));
  for (int a = 0; a < n_nodes; ++a) {
    int cnt = aptr[a + 1] - aptr[a];
    qsort(adj + aptr[a], (size_t)cnt, sizeof(int), cmp_int);
    int u = 0;
    for (int k = 0; k < cnt; ++k)
      if (u == 0 || adj[aptr[a] + k] != nadj[nptr[a] + u - 1]) nadj[nptr[a] + u++] = adj[aptr[a] + k];
    if (cnt == 0) nadj[nptr[a] + u++] = a; /* isolated node keeps its diagonal */
    nptr[a + 1] = nptr[a] + u;
  }
  free(deg); free(aptr); free(adj); free(fill);
  /* scalar CSR: row 3a+r has columns 3b+c for b in nadj(a), c = 0..2 */
  int n = 3 * n_nodes;
  int *ptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
  for (int a = 0; a < n_nodes; ++a)
    for (int r = 0; r < 3; ++r) ptr[3 * a + r + 1] = 3 * (nptr[a + 1] - nptr[a]);
  for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  int *idx = (int *)xcalloc((size_t)ptr[n], sizeof(int));
  double *val = (double *)xcalloc((size_t)ptr[n], sizeof(double));
  double *kval = (double *)xcalloc((size_t)ptr[n], sizeof(double)); /* unscaled K */
  for (int a = 0; a < n_nodes; ++a)
    for (int r = 0; r < 3; ++r) {
      int o = ptr[3 * a + r];
      for (int k = nptr[a]; k < nptr[a + 1]; ++k)
        for (int c = 0; c < 3; ++c) idx[o++] = 3 * nadj[k] + c;
    }
  for (int i = 0; i < n; ++i) b[i] = 0.0;
  /* materialEnergyEval: per tet, in tet order (FEMEnergy.cpp:241-250, 98-106);
     stiffness triplets are scaled by KCoef + CCoef beta = h^2 + h alpha_m
     (FEMSystem.cpp:180-181) before summation.  buildLinearSystem's
     coefficients with beta = 1 (MprgpFemSolver.cpp:331-339): bdt = h,
     KCoef = FCoef = bdt h, CCoef = bdt, MCoef = 1. */
  double bdt = 1.0 * h;
  double h2 = bdt * h;
  double kc = h2 + bdt * alpha_m;
  for (int t = 0; t < n_tets; ++t) {
    const int *v = tets + 4 * t;
    double xe[12], fB[12], HB[144];
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r) xe[i * 3 + r] = x1[3 * v[i] + r];
    double E = young[t], nu = poisson[t];
    double mu = E / (2.0 * (1.0 + nu));
    double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    orc_fem_element(type, xe, d + 9 * t, vol[t], lam, mu, fB, HB);
    for (int i = 0; i < 4; ++i) {
      for (int r = 0; r < 3; ++r) b[3 * v[i] + r] += fB[3 * i + r] * h2;
      for (int j = 0; j < 4; ++j) {
        /* locate node v[j] in the adjacency of v[i] */
        int lo = nptr[v[i]], hi = nptr[v[i] + 1] - 1;
        while (lo < hi) {
          int mid = (lo + hi) / 2;
          if (nadj[mid] < v[j]) lo = mid + 1; else hi = mid;
        }
        int kb = lo - nptr[v[i]];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            val[ptr[3 * v[i] + r] + 3 * kb + c] += HB[(3 * i + r) * 12 + 3 * j + c] * kc;
            kval[ptr[3 * v[i] + r] + 3 * kb + c] += HB[(3 * i + r) * 12 + 3 * j + c];
          }
      }
    }
  }
  /* RHS += (K CRHS) beta, CRHS = (x0 - x1) bdt with x0 = x_n (FEMSystem.cpp:179) */
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += kval[k] * ((x_n[idx[k]] - x1[idx[k]]) * bdt);
    b[i] += s * alpha_m;
  }
  free(kval);
  /* gravity (FixedForceEnergy, FEMEnergy.cpp:108-115), mass term and diagonal */
  for (int a = 0; a < n_nodes; ++a) {
    double M = node_mass[a] * density;
    for (int r = 0; r < 3; ++r) {
      b[3 * a + r] += ((gravity[r] * density) * node_mass[a]) * h2;
    }
    for (int r = 0; r < 3; ++r) {
      double psi = x_n[3 * a + r] + h * v_n[3 * a + r];
      double crhs = (x_n[3 * a + r] - x1[3 * a + r]) * bdt;
      b[3 * a + r] += M * ((psi - x1[3 * a + r]) + crhs * alpha_k);
      /* diagonal entry of row 3a+r */
      int lo = nptr[a], hi = nptr[a + 1] - 1;
      while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (nadj[mid] < a) lo = mid + 1; else hi = mid;
      }
      int kb = lo - nptr[a];
      val[ptr[3 * a + r] + 3 * kb + r] += M * (1.0 + bdt * alpha_k);
    }
  }
  /* RHS += LHS * x1 (MprgpFemSolver.cpp:349) */
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x1[idx[k]];
    b[i] += s;
  }
  free(nptr); free(nadj);
  *ptr_out = ptr; *idx_out = idx; *val_out = val;
  return 0;
}

int orc_fem_build_system_at(int n_nodes, int n_tets, const int *tets,
                            const double *d, const double *vol, const double *node_mass,
                            const double *young, const double *poisson, double density,
                            const double *x_n, const double *v_n, const double *x1,
                            const double *gravity, double h, int type,
                            int **ptr_out, int **idx_out, double **val_out, double *b) {
  return orc_fem_build_system_damped(n_nodes, n_tets, tets, d, vol, node_mass, young, poisson, density,
                                     x_n, v_n, x1, gravity, h, type, 0.0, 0.0, ptr_out, idx_out,
                                     val_out, b);
}

int orc_fem_build_system(int n_nodes, int n_tets, const int *tets,
                         const double *d, const double *vol, const double *node_mass,
                         const double *young, const double *poisson, double density,
                         const double *x_n, const double *v_n, const double *gravity,
                         double h, int type,
                         int **ptr_out, int **idx_out, double **val_out, double *b) {
  return orc_fem_build_system_at(n_nodes, n_tets, tets, d, vol, node_mass, young, poisson, density, x_n,
                                 v_n, x_n, gravity, h, type, ptr_out, idx_out, val_out, b);
}

/* updateMesh under implicit Euler (beta = 1, beta2 = 0.5, gamma = 1): the
 * previous acceleration enters PHI and PSI with the factors 1 - gamma = 0 and
 * 1 - 2 beta2 = 0, so a = (x - (x_n + h v_n)) / (h h), v = v_n + a h. */
void orc_implicit_euler_velocity(int n, const double *x_n, const double *v_n, const double *x,
                                 double h, double *v) {
  for (int i = 0; i < n; ++i) {
    double psi = x_n[i] + h * v_n[i];
    double a = (x[i] - psi) / ((1.0 * h) * h);
    v[i] = v_n[i] + (a * 1.0) * h;
  }
}
