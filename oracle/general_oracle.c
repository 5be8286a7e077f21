/*
 * general_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Serial FP64 restatement of the reference's coupled-constraint MPRGP solve:
 * MPRGPGeneralCon<double>::solve (MPRGPSolver.h:594-626) with
 * GeneralConProjector (MPRGPProjection.h:400-520), its base
 * BaseGeneralConProjector (:241-317) and DiagonalInFacePreconSolver
 * (MPRGPPrecondition.h:36-69), over MPRGPMonotonic::solve
 * (MPRGPSolver.h:148-316, 412-454).  Rows J x >= c may span several nodes
 * (self-contact rows of 12 nonzeros, LinearConCollider.cpp:186-217).
 *
 * The reference solves its small dense QPs with QuadProg++ (Goldfarb-Idnani,
 * SimpleQP.h:16-48) and the active rows' normal equations with a sparse LLT.
 * Both have unique answers for a J of full row rank, so the oracle computes
 * the same quantities by its own exact methods: orc_qp_nonneg is a primal
 * active-set method on the free set's Cholesky factor, and PHI solves the
 * active rows' normal equations by dense Cholesky.
 *
 * Function -> reference -> pin (tests/test_oracle_general.py):
 *   orc_qp_nonneg       solveQP(A, b, I, 0) (SimpleQP.h:16-48)     brute-force enumeration of active sets, KKT
 *   orc_gproject        GeneralConProjector::project :412-431      brute-force nearest feasible point; reduces to
 *                                                                  the decoupled projection for decoupled J
 *   orc_gphi            PHI :432-451                               J_a phi = 0, g - phi in the row space of J_a
 *   orc_gbeta           BETA :452-473                              KKT of its QP; reduces to the decoupled BETA
 *   orc_gphitphi        PHITPHI :474-484                           closed form with orc_gproject
 *   orc_general_solve   MPRGPGeneralCon::solve :594-626            brute-force minimiser of tiny coupled QPs, KKT
 *                                                                  on a beam with self-contact rows
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mprgp_oracle.h"

static const double G_EPS = 1e-9; /* ScalarUtil<double>::scalar_eps (Simulator.cpp:17) */

static void *gcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

static double gdot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* In-place dense Cholesky of the k x k leading block of M (row-major, lda):
 * returns 0 on success, -1 if a pivot is not positive. */
static int chol(double *M, int k, int lda) {
  for (int j = 0; j < k; ++j) {
    double d = M[j * lda + j];
    for (int p = 0; p < j; ++p) d -= M[j * lda + p] * M[j * lda + p];
    if (!(d > 0.0)) return -1;
    d = sqrt(d);
    M[j * lda + j] = d;
    for (int i = j + 1; i < k; ++i) {
      double s = M[i * lda + j];
      for (int p = 0; p < j; ++p) s -= M[i * lda + p] * M[j * lda + p];
      M[i * lda + j] = s / d;
    }
  }
  return 0;
}

/* Solves (L L^T) x = b in place (L from chol). */
static void chol_solve(const double *L, int k, int lda, double *b) {
  for (int i = 0; i < k; ++i) {
    double s = b[i];
    for (int p = 0; p < i; ++p) s -= L[i * lda + p] * b[p];
    b[i] = s / L[i * lda + i];
  }
  for (int i = k - 1; i >= 0; --i) {
    double s = b[i];
    for (int p = i + 1; p < k; ++p) s -= L[p * lda + i] * b[p];
    b[i] = s / L[i * lda + i];
  }
}

/* lam = argmin 1/2 l'Gl - h'l subject to l >= 0, G (m x m, row-major) SPD:
 * a primal active-set method.  The free set F starts empty; each outer step
 * frees the bound variable with the largest positive w_i = (h - G l)_i, then
 * solves G_FF s = h_F; while some free s_i <= 0 it steps from l towards s up
 * to the first zero and bounds that variable.  Returns 0, or -1 if G is not
 * positive definite on a free set or the iteration limit is reached. */
int orc_qp_nonneg(int m, const double *G, const double *h, double *lam) {
  int *free_ = (int *)gcalloc((size_t)m, sizeof(int));
  int *idx = (int *)gcalloc((size_t)m, sizeof(int));
  double *L = (double *)gcalloc((size_t)m * (size_t)m, sizeof(double));
  double *s = (double *)gcalloc((size_t)m, sizeof(double));
  double *w = (double *)gcalloc((size_t)m, sizeof(double));
  int rc = -1;
  for (int i = 0; i < m; ++i) lam[i] = 0.0;
  double hscale = 0.0;
  for (int i = 0; i < m; ++i) hscale = fmax(hscale, fabs(h[i]));
  const double tol = 1e-14 * (hscale > 0.0 ? hscale : 1.0);
  for (int outer = 0; outer < 10 * m + 10; ++outer) {
    /* w = h - G lam; stop when no bound variable wants to grow */
    int best = -1;
    double wbest = tol;
    for (int i = 0; i < m; ++i) {
      double r = h[i];
      for (int j = 0; j < m; ++j) r -= G[i * m + j] * lam[j];
      w[i] = r;
      if (!free_[i] && r > wbest) { wbest = r; best = i; }
    }
    if (best < 0) { rc = 0; break; }
    free_[best] = 1;
    for (int inner = 0; inner <= m; ++inner) {
      int k = 0;
      for (int i = 0; i < m; ++i) if (free_[i]) idx[k++] = i;
      for (int a = 0; a < k; ++a) {
        for (int b = 0; b < k; ++b) L[a * m + b] = G[idx[a] * m + idx[b]];
        s[a] = h[idx[a]];
      }
      if (chol(L, k, m)) goto done;
      chol_solve(L, k, m, s);
      double step = 1.0;
      int hit = -1;
      for (int a = 0; a < k; ++a)
        if (s[a] <= 0.0) {
          const double t = lam[idx[a]] / (lam[idx[a]] - s[a]);
          if (t < step) { step = t; hit = a; }
        }
      if (hit < 0) {
        for (int i = 0; i < m; ++i) lam[i] = 0.0;
        for (int a = 0; a < k; ++a) lam[idx[a]] = s[a];
        break;
      }
      for (int a = 0; a < k; ++a) lam[idx[a]] += step * (s[a] - lam[idx[a]]);
      free_[idx[hit]] = 0;
      lam[idx[hit]] = 0.0;
      for (int a = 0; a < k; ++a)
        if (lam[idx[a]] <= 0.0) { free_[idx[a]] = 0; lam[idx[a]] = 0.0; }
    }
  }
done:
  free(free_); free(idx); free(L); free(s); free(w);
  return rc;
}

/* ------------------------------------------------- GeneralConProjector */

struct orc_gprojector {
  int m, n;
  const int *ptr, *idx;
  const double *val, *c;
  double *JJt;      /* dense m x m */
  char *face;
  int n_active;
  int *act;         /* active rows, in order */
  double *La;       /* Cholesky factor of J_a J_a^T (n_active x n_active, lda m) */
  double *t_m, *t_m2, *t_n;
};

static void g_jmul(const orc_gprojector *P, const double *x, double *out) {
  for (int i = 0; i < P->m; ++i) {
    double s = 0.0;
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) s += P->val[k] * x[P->idx[k]];
    out[i] = s;
  }
}

/* out = J^T lam (rows in ascending order). */
static void g_jtmul(const orc_gprojector *P, const double *lam, double *out) {
  memset(out, 0, sizeof(double) * (size_t)P->n);
  for (int i = 0; i < P->m; ++i)
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) out[P->idx[k]] += P->val[k] * lam[i];
}

orc_gprojector *orc_gprojector_new(const orc_csr *J, const double *c) {
  orc_gprojector *P = (orc_gprojector *)gcalloc(1, sizeof(orc_gprojector));
  P->m = J->rows;
  P->n = J->cols;
  P->ptr = J->ptr;
  P->idx = J->idx;
  P->val = J->val;
  P->c = c;
  const int m = P->m;
  P->JJt = (double *)gcalloc((size_t)m * (size_t)m, sizeof(double));
  P->face = (char *)gcalloc((size_t)m, 1);
  P->act = (int *)gcalloc((size_t)m, sizeof(int));
  P->La = (double *)gcalloc((size_t)m * (size_t)m, sizeof(double));
  P->t_m = (double *)gcalloc((size_t)m, sizeof(double));
  P->t_m2 = (double *)gcalloc((size_t)m, sizeof(double));
  P->t_n = (double *)gcalloc((size_t)P->n, sizeof(double));
  /* JJt = J J^T (the constructor, :409-411) by a column-dense pass per row */
  double *row = (double *)gcalloc((size_t)P->n, sizeof(double));
  for (int i = 0; i < m; ++i) {
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) row[P->idx[k]] += P->val[k];
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int k = P->ptr[j]; k < P->ptr[j + 1]; ++k) s += P->val[k] * row[P->idx[k]];
      P->JJt[i * m + j] = s;
    }
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) row[P->idx[k]] = 0.0;
  }
  free(row);
  return P;
}

void orc_gprojector_free(orc_gprojector *P) {
  if (!P) return;
  free(P->JJt); free(P->face); free(P->act); free(P->La); free(P->t_m); free(P->t_m2); free(P->t_n);
  free(P);
}

const char *orc_gface(const orc_gprojector *P) { return P->face; }
int orc_gnum_active(const orc_gprojector *P) { return P->n_active; }

/* GeneralConProjector::project (:412-431): y = x + J^T lam with
 * lam = argmin 1/2 l'(J J^T)l - (c - J x)'l, l >= 0.  Returns 0, or -1 when
 * the QP fails (J J^T singular on the rows it frees). */
int orc_gproject(const orc_gprojector *P, const double *x, double *y) {
  if (P->m == 0) {
    if (y != x) memcpy(y, x, sizeof(double) * (size_t)P->n);
    return 0;
  }
  double *h = P->t_m, *lam = P->t_m2;
  g_jmul(P, x, h);
  for (int i = 0; i < P->m; ++i) h[i] = P->c[i] - h[i];
  const int rc = orc_qp_nonneg(P->m, P->JJt, h, lam);
  g_jtmul(P, lam, P->t_n);
  for (int j = 0; j < P->n; ++j) y[j] = x[j] + P->t_n[j];
  return rc;
}

/* DECIDE_FACE (:485-489 over BaseGeneralConProjector :283-297) and
 * prepareActiveConMatrix (:495-520): the active rows and the Cholesky
 * factor of J_a J_a^T (SimplicialLLT in PHI, :440-441). */
int orc_gdecide_face(orc_gprojector *P, const double *x) {
  double *jx = P->t_m;
  g_jmul(P, x, jx);
  P->n_active = 0;
  for (int i = 0; i < P->m; ++i) {
    P->face[i] = (fabs(jx[i] - P->c[i]) < G_EPS) ? 1 : 0;
    if (P->face[i]) P->act[P->n_active++] = i;
  }
  const int k = P->n_active, m = P->m;
  for (int a = 0; a < k; ++a)
    for (int b = 0; b < k; ++b) P->La[a * m + b] = P->JJt[P->act[a] * m + P->act[b]];
  return chol(P->La, k, m);
}

int orc_gis_feasible(const orc_gprojector *P, const double *x) {
  double *jx = P->t_m;
  g_jmul(P, x, jx);
  for (int i = 0; i < P->m; ++i)
    if (jx[i] < P->c[i] - G_EPS) return 0;
  return 1;
}

/* PHI (:432-451): phi = g + J_a^T lam_a with (J_a J_a^T) lam_a = -J_a g. */
void orc_gphi(const orc_gprojector *P, const double *g, double *phi) {
  if (P->n_active == 0) {
    if (phi != g) memcpy(phi, g, sizeof(double) * (size_t)P->n);
    return;
  }
  double *jg = P->t_m, *lam = P->t_m2;
  g_jmul(P, g, jg);
  const int k = P->n_active;
  double *la = (double *)gcalloc((size_t)k, sizeof(double));
  for (int a = 0; a < k; ++a) la[a] = -jg[P->act[a]];
  chol_solve(P->La, k, P->m, la);
  for (int i = 0; i < P->m; ++i) lam[i] = 0.0;
  for (int a = 0; a < k; ++a) lam[P->act[a]] = la[a];
  free(la);
  g_jtmul(P, lam, P->t_n);
  for (int j = 0; j < P->n; ++j) phi[j] = g[j] + P->t_n[j];
}

/* BETA (:452-473): gp = g - phi, beta = gp - J_a^T lam with
 * lam = argmin 1/2 l'(J_a J_a^T) l - (J_a gp)'l, l >= 0. */
int orc_gbeta(const orc_gprojector *P, const double *g, const double *phi, double *beta) {
  for (int j = 0; j < P->n; ++j) beta[j] = g[j] - phi[j];
  if (P->n_active == 0) return 0;
  const int k = P->n_active, m = P->m;
  double *jg = P->t_m, *lam = P->t_m2;
  g_jmul(P, beta, jg);
  double *Ga = (double *)gcalloc((size_t)k * (size_t)k, sizeof(double));
  double *h = (double *)gcalloc((size_t)k, sizeof(double));
  double *la = (double *)gcalloc((size_t)k, sizeof(double));
  for (int a = 0; a < k; ++a) {
    h[a] = jg[P->act[a]];
    for (int b = 0; b < k; ++b) Ga[a * k + b] = P->JJt[P->act[a] * m + P->act[b]];
  }
  const int rc = orc_qp_nonneg(k, Ga, h, la);
  for (int i = 0; i < m; ++i) lam[i] = 0.0;
  for (int a = 0; a < k; ++a) lam[P->act[a]] = la[a];
  free(Ga); free(h); free(la);
  g_jtmul(P, lam, P->t_n);
  for (int j = 0; j < P->n; ++j) beta[j] -= P->t_n[j];
  return rc;
}

/* PHITPHI (:474-484). */
double orc_gphitphi(const orc_gprojector *P, const double *x, double alpha_bar, const double *phi) {
  const int n = P->n;
  double *xa = (double *)gcalloc((size_t)n, sizeof(double));
  double *px = (double *)gcalloc((size_t)n, sizeof(double));
  for (int j = 0; j < n; ++j) xa[j] = x[j] - alpha_bar * phi[j];
  orc_gproject(P, xa, px);
  for (int j = 0; j < n; ++j) xa[j] = x[j] - px[j];
  const double r = gdot(n, xa, phi) * (1.0 / alpha_bar);
  free(xa); free(px);
  return r > 0.0 ? r : 0.0;
}

/* ------------------------------------------------ MPRGPMonotonic, general */

typedef struct {
  const orc_csr *A;
  const double *B;
  orc_gprojector *P;
  const double *invD; /* DiagonalInFacePreconSolver: 1/diag(A), or NULL (no preconditioning) */
  int n, max_it, iteration, num_cg, num_exp, num_prop, y_size, qp_fail;
  double tol, alpha_bar, residual;
  double *g, *p, *z, *beta, *phi, *gp, *y, *t;
} gsolver;

static void g_grad(gsolver *S, const double *x) {
  orc_spmv(S->A, x, S->g);
  for (int i = 0; i < S->n; ++i) S->g[i] -= S->B[i];
}

/* DiagonalInFacePreconSolver::solve (MPRGPPrecondition.h:53-64): z = phi D^-1. */
static void g_precond(gsolver *S) {
  if (S->invD)
    for (int i = 0; i < S->n; ++i) S->z[i] = S->phi[i] * S->invD[i];
  else
    memcpy(S->z, S->phi, sizeof(double) * (size_t)S->n);
}

static void g_face(gsolver *S, const double *x) {
  if (orc_gdecide_face(S->P, x)) S->qp_fail = 1;
}

static double g_gradients(gsolver *S, int comp_phi) {
  if (comp_phi) orc_gphi(S->P, S->g, S->phi);
  if (orc_gbeta(S->P, S->g, S->phi, S->beta)) S->qp_fail = 1;
  for (int i = 0; i < S->n; ++i) S->gp[i] = S->phi[i] + S->beta[i];
  return sqrt(gdot(S->n, S->gp, S->gp));
}

static int g_proportional(gsolver *S, const double *x) {
  return gdot(S->n, S->beta, S->beta) <= orc_gphitphi(S->P, x, S->alpha_bar, S->phi);
}

static double g_fproj(gsolver *S, const double *x) {
  if (orc_gproject(S->P, x, S->t)) S->qp_fail = 1;
  return orc_func_value(S->A, S->t, S->B);
}

static void g_exp_step(gsolver *S, double *x) {
  S->num_exp++;
  if (orc_gproject(S->P, x, S->y)) S->qp_fail = 1;
  S->y_size = S->n;
  g_grad(S, S->y);
  g_face(S, S->y);
  orc_gphi(S->P, S->g, S->phi);
  for (int i = 0; i < S->n; ++i) S->y[i] -= S->alpha_bar * S->phi[i];
  if (orc_gproject(S->P, S->y, x)) S->qp_fail = 1;
  g_grad(S, x);
  g_face(S, x);
  orc_gphi(S->P, S->g, S->phi);
  g_precond(S);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
}

static void g_prop_step(gsolver *S, double *x) {
  S->num_prop++;
  double *D = S->beta, *AD = S->gp;
  orc_spmv(S->A, D, AD);
  const double alpha = gdot(S->n, S->g, D) / gdot(S->n, D, AD);
  for (int i = 0; i < S->n; ++i) x[i] -= alpha * D[i];
  memcpy(S->t, x, sizeof(double) * (size_t)S->n);
  if (orc_gproject(S->P, S->t, x)) S->qp_fail = 1;
  for (int i = 0; i < S->n; ++i) S->g[i] -= alpha * AD[i];
  g_face(S, x);
  orc_gphi(S->P, S->g, S->phi);
  g_precond(S);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
}

static void g_cg_steps(gsolver *S, double *x, int *prop) {
  const int n = S->n;
  double *AP = S->gp;
  orc_spmv(S->A, S->p, AP);
  double pd = gdot(n, S->p, AP);
  double alpha = gdot(n, S->z, S->g) / pd;
  for (int i = 0; i < n; ++i) S->y[i] = x[i] - alpha * S->p[i];
  S->y_size = n;
  double fy = g_fproj(S, S->y);
  double fx = orc_func_value(S->A, x, S->B);
  *prop = 1;
  while (*prop && S->residual > S->tol && fy <= fx && S->iteration < S->max_it) {
    memcpy(x, S->y, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) S->g[i] -= alpha * AP[i];
    orc_gphi(S->P, S->g, S->phi);
    g_precond(S);
    const double beta = gdot(n, S->z, AP) / gdot(n, S->p, AP);
    for (int i = 0; i < n; ++i) S->p[i] = S->z[i] - beta * S->p[i];
    S->residual = g_gradients(S, 0);
    *prop = g_proportional(S, x);
    orc_spmv(S->A, S->p, AP);
    pd = gdot(n, S->p, AP);
    alpha = gdot(n, S->z, S->g) / pd;
    for (int i = 0; i < n; ++i) S->y[i] = x[i] - alpha * S->p[i];
    fx = fy;
    fy = g_fproj(S, S->y);
    S->num_cg++;
    S->iteration++;
  }
}

/* MPRGPGeneralCon<double>::solve (MPRGPSolver.h:611-623, J/c overload):
 * projects x, then MPRGPMonotonic::solve (:412-454) with the general
 * projector and DiagonalInFacePreconSolver<!preconditioned>.  alpha_bar > 0
 * is used as is, else 2/specRad(A, ev, tol).  Returns 0 converged, -1 not,
 * -2 when a projection QP or an active-row factorisation failed. */
int orc_general_solve(const orc_csr *A, const double *B, const orc_csr *J, const double *c, double *x, double tol,
                      int max_it, int preconditioned, double alpha_bar, double *ev, int *ev_valid,
                      int project_first, orc_stats *st) {
  const int n = A->rows;
  gsolver S;
  memset(&S, 0, sizeof(S));
  S.A = A; S.B = B; S.n = n;
  S.P = orc_gprojector_new(J, c);
  if (project_first) {
    double *x0 = (double *)gcalloc((size_t)n, sizeof(double));
    memcpy(x0, x, sizeof(double) * (size_t)n);
    if (orc_gproject(S.P, x0, x)) S.qp_fail = 1;
    free(x0);
  }
  double *invD = NULL;
  if (preconditioned) {
    invD = (double *)gcalloc((size_t)n, sizeof(double));
    orc_get_diagonal(A, invD);
    for (int i = 0; i < n; ++i) invD[i] = 1.0 / invD[i];
  }
  S.invD = invD;
  S.max_it = max_it;
  S.tol = tol < 1e-30 ? 1e-30 : tol;
  int piters = 0;
  double lmax;
  if (alpha_bar > 0.0) {
    S.alpha_bar = alpha_bar;
    lmax = 2.0 / alpha_bar;
  } else {
    lmax = orc_spec_rad(A, ev, ev_valid, tol, 0x5EEDULL, &piters);
    S.alpha_bar = 2.0 / lmax;
  }
  double **vecs[] = {&S.g, &S.p, &S.z, &S.beta, &S.phi, &S.gp, &S.y, &S.t};
  for (int v = 0; v < 8; ++v) *vecs[v] = (double *)gcalloc((size_t)n, sizeof(double));
  /* initialize (:148-162) */
  g_grad(&S, x);
  g_face(&S, x);
  orc_gphi(S.P, S.g, S.phi);
  g_precond(&S);
  memcpy(S.p, S.z, sizeof(double) * (size_t)n);
  S.residual = 0.0;
  for (S.iteration = 0; S.iteration < S.max_it;) {
    S.residual = g_gradients(&S, 1);
    if (S.residual <= S.tol) break;
    int prop = g_proportional(&S, x);
    if (prop) g_cg_steps(&S, x, &prop);
    if (S.residual <= S.tol) break;
    if (prop || S.y_size <= 0 || !orc_gis_feasible(S.P, S.y)) {
      g_exp_step(&S, x);
    } else {
      g_grad(&S, x);
      g_prop_step(&S, x);
    }
    S.iteration++;
  }
  memcpy(S.t, x, sizeof(double) * (size_t)n);
  if (orc_gproject(S.P, S.t, x)) S.qp_fail = 1;
  const int code = S.qp_fail ? -2 : (S.residual <= S.tol ? 0 : -1);
  if (st) {
    st->code = code;
    st->iterations = S.iteration;
    st->num_cg = S.num_cg;
    st->num_exp = S.num_exp;
    st->num_prop = S.num_prop;
    st->active = S.P->n_active;
    st->power_iters = piters;
    st->residual = S.residual;
    st->alpha_bar = S.alpha_bar;
    st->lambda_max = lmax;
  }
  for (int v = 0; v < 8; ++v) free(*vecs[v]);
  free(invD);
  orc_gprojector_free(S.P);
  return code;
}
