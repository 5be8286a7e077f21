/*
 * mprgp_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Serial FP64 restatement of the reference's decoupled MPRGP solve.  Built
 * with -ffp-contract=off so every operation rounds once, as the reference's
 * SSE2 Eigen build does.  Sums run sequentially (Eigen's packet reductions
 * are not reproduced), so the oracle agrees with the reference in step logic
 * and to rounding in values; bit-level agreement with an Eigen build is not
 * claimed.
 *
 * Function -> reference (file:line) -> how it is pinned (tests/test_oracle_solver.py):
 *   orc_spmv            MPRGPSolver.h:469-473      dense numpy matvec on random sparse SPD
 *   orc_func_value      MPRGPSolver.h:477-482      closed form 0.5 x'Ax - x'b in numpy
 *   orc_project         MPRGPProjection.h:333-348  brute-force nearest feasible point (per row half-space), idempotence
 *   orc_decide_face     MPRGPProjection.h:283-297  |Jx-c|<1e-9 by numpy
 *   orc_is_feasible     MPRGPProjection.h:298-310  Jx >= c-1e-9 by numpy
 *   orc_phi / orc_beta  MPRGPProjection.h:349-383  tangent projection invariants (n.phi=0 on active rows,
 *                                                  phi+beta = g on free rows, beta = n min(0,n.g) on active)
 *   orc_phitphi         MPRGPProjection.h:384-393  closed form on hand-built cases
 *   orc_step_limit      MPRGPProjection.h:263-282  brute-force largest feasible step on random data
 *   orc_precond_*       MPRGPPrecondition.h:36-110 n.z = 0 on active rows, z = D^-1 g on free nodes,
 *                                                  z = phi with no rows (unpreconditioned quirk)
 *   orc_spec_rad        MPRGPSolver.h:42-79        numpy eigvalsh lambda_max on small SPD matrices
 *   orc_decoupled_solve MPRGPSolver.h:148-456, 538-574
 *                                                  exact QP minimiser by brute-force active-set enumeration
 *                                                  (<= 8 nodes), KKT conditions on beam QPs, the m = 0 case
 *                                                  against a dense Cholesky solve, the reference's own
 *                                                  test_solver case (J = I, n = 5) returning code 0
 *   orc_decoupled_solve_ex  the same solve with its decisions recorded and an optional closed-form
 *                       f(P(y)) (the device mixed form's): tracing changes no bit of the solve (identical
 *                       x and counts with and without it), the recorded step counts equal the stats, the
 *                       closed form equals the direct f(P(y)) to 1e-13 relative along trajectories that
 *                       accept infeasible trial points, and a planted wrong continuation (from f(P(y)))
 *                       is caught (tests/test_oracle_trace.py)
 */
#include "mprgp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ScalarUtil<double> values used by the simulator (Simulator.cpp:17-18). */
static const double ORC_EPS = 1e-9;
static const double ORC_MAX = DBL_MAX;

/* ---------------------------------------------------------------- helpers */

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static double norm2(int n, const double *a) { return sqrt(dot(n, a, a)); }

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

void orc_free(void *p) { free(p); }

/* FixedSparseMatrix::multiply (MPRGPSolver.h:469-473): y = A x. */
void orc_spmv(const orc_csr *A, const double *x, double *y) {
  for (int i = 0; i < A->rows; ++i) {
    double s = 0.0;
    for (int k = A->ptr[i]; k < A->ptr[i + 1]; ++k) s += A->val[k] * x[A->idx[k]];
    y[i] = s;
  }
}

/* FixedSparseMatrix::funcValue (MPRGPSolver.h:477-482): 0.5 x'Ax - x'b. */
double orc_func_value(const orc_csr *A, const double *x, const double *b) {
  double *ax = (double *)xcalloc((size_t)A->rows, sizeof(double));
  orc_spmv(A, x, ax);
  double f = 0.5 * dot(A->rows, x, ax) - dot(A->rows, x, b);
  free(ax);
  return f;
}

/* getDiagonal (MPRGPUtility.h:503-515). */
void orc_get_diagonal(const orc_csr *A, double *diag) {
  for (int i = 0; i < A->rows; ++i) {
    diag[i] = 0.0;
    for (int k = A->ptr[i]; k < A->ptr[i + 1]; ++k)
      if (A->idx[k] == i) { diag[i] = A->val[k]; break; }
  }
}

/* ------------------------------------------------------------- projector */

struct orc_projector {
  int m, n;
  const int *ptr, *idx;
  const double *val;
  const double *c;
  double *JJt;      /* diag(J J^T)                       */
  char *face;
  int num_active;
  double *lam;      /* m scratch                          */
  double *tn;       /* n scratch for J^T lambda           */
  double *tn2;      /* n scratch                          */
};

static void j_mul(const orc_projector *P, const double *x, double *out) {
  for (int i = 0; i < P->m; ++i) {
    double s = 0.0;
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) s += P->val[k] * x[P->idx[k]];
    out[i] = s;
  }
}

/* out = J^T lam, accumulating each column over rows in ascending order. */
static void jt_mul(const orc_projector *P, const double *lam, double *out) {
  memset(out, 0, sizeof(double) * (size_t)P->n);
  for (int i = 0; i < P->m; ++i)
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) out[P->idx[k]] += P->val[k] * lam[i];
}

orc_projector *orc_projector_new(const orc_csr *J, const double *c) {
  orc_projector *P = (orc_projector *)xcalloc(1, sizeof(orc_projector));
  P->m = J->rows;
  P->n = J->cols;
  P->ptr = J->ptr;
  P->idx = J->idx;
  P->val = J->val;
  P->c = c;
  P->JJt = (double *)xcalloc((size_t)P->m, sizeof(double));
  P->face = (char *)xcalloc((size_t)P->m, 1);
  P->lam = (double *)xcalloc((size_t)P->m, sizeof(double));
  P->tn = (double *)xcalloc((size_t)P->n, sizeof(double));
  P->tn2 = (double *)xcalloc((size_t)P->n, sizeof(double));
  /* MPRGPSolver.h:566-569: JJt = diag(J J^T). */
  for (int i = 0; i < P->m; ++i) {
    double s = 0.0;
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) s += P->val[k] * P->val[k];
    P->JJt[i] = s;
  }
  P->num_active = 0;
  return P;
}

void orc_projector_free(orc_projector *P) {
  if (!P) return;
  free(P->JJt); free(P->face); free(P->lam); free(P->tn); free(P->tn2);
  free(P);
}

/* The decoupled facade asserts that J J^T is structurally diagonal
 * (MPRGPSolver.h:566-567): no two rows share a column. */
int orc_projector_is_decoupled(const orc_projector *P) {
  int *owner = (int *)xcalloc((size_t)P->n, sizeof(int));
  int ok = 1;
  for (int i = 0; i < P->m && ok; ++i)
    for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) {
      int col = P->idx[k];
      if (owner[col] && owner[col] != i + 1) { ok = 0; break; }
      owner[col] = i + 1;
    }
  free(owner);
  return ok;
}

const char *orc_face(const orc_projector *P) { return P->face; }
int orc_num_active(const orc_projector *P) { return P->num_active; }
void orc_set_face(orc_projector *P, const char *face) {
  P->num_active = 0;
  for (int i = 0; i < P->m; ++i) { P->face[i] = face[i]; P->num_active += face[i] != 0; }
}

/* DecoupledConProjector::project (MPRGPProjection.h:333-348). */
void orc_project(const orc_projector *P, const double *x, double *y) {
  if (P->m == 0) {
    if (y != x) memcpy(y, x, sizeof(double) * (size_t)P->n);
    return;
  }
  double *lam = P->lam;
  j_mul(P, x, lam);
  for (int i = 0; i < P->m; ++i) {
    double v = (P->c[i] - lam[i]) / P->JJt[i];
    lam[i] = (v < 0.0) ? 0.0 : v; /* std::max<T>(v, 0.0) */
  }
  jt_mul(P, lam, P->tn);
  for (int j = 0; j < P->n; ++j) y[j] = x[j] + P->tn[j];
}

/* BaseGeneralConProjector::DECIDE_FACE (MPRGPProjection.h:283-297). */
void orc_decide_face(orc_projector *P, const double *x) {
  j_mul(P, x, P->lam);
  P->num_active = 0;
  for (int i = 0; i < P->m; ++i) {
    P->face[i] = (fabs(P->lam[i] - P->c[i]) < ORC_EPS) ? 1 : 0;
    P->num_active += P->face[i];
  }
}

/* BaseGeneralConProjector::isFeasible (MPRGPProjection.h:298-310). */
int orc_is_feasible(const orc_projector *P, const double *x) {
  j_mul(P, x, P->lam);
  for (int i = 0; i < P->m; ++i)
    if (P->lam[i] < P->c[i] - ORC_EPS) return 0;
  return 1;
}

/* DecoupledConProjector::PHI (MPRGPProjection.h:349-367). */
void orc_phi(const orc_projector *P, const double *g, double *phi) {
  if (P->num_active == 0) {
    if (phi != g) memcpy(phi, g, sizeof(double) * (size_t)P->n);
    return;
  }
  double *lam = P->lam;
  j_mul(P, g, lam);
  for (int i = 0; i < P->m; ++i) lam[i] = (P->face[i] == 0) ? 0.0 : -lam[i] / P->JJt[i];
  jt_mul(P, lam, P->tn);
  for (int j = 0; j < P->n; ++j) phi[j] = g[j] + P->tn[j];
}

/* DecoupledConProjector::BETA (MPRGPProjection.h:368-383).  Rows that are
 * not in the face keep lambda = (J beta)_i unscaled, as the reference does. */
void orc_beta(const orc_projector *P, const double *g, const double *phi, double *beta) {
  for (int j = 0; j < P->n; ++j) beta[j] = g[j] - phi[j];
  if (P->num_active == 0) return;
  double *lam = P->lam;
  j_mul(P, beta, lam);
  for (int i = 0; i < P->m; ++i)
    if (P->face[i] != 0) {
      double v = lam[i] / P->JJt[i];
      lam[i] = (0.0 < v) ? v : 0.0; /* std::max<T>(0.0, v) */
    }
  jt_mul(P, lam, P->tn);
  for (int j = 0; j < P->n; ++j) beta[j] -= P->tn[j];
}

/* DecoupledConProjector::PHITPHI (MPRGPProjection.h:384-393). */
double orc_phitphi(const orc_projector *P, const double *x, double alpha_bar, const double *phi) {
  int n = P->n;
  double *xa = (double *)xcalloc((size_t)n, sizeof(double));
  double *px = (double *)xcalloc((size_t)n, sizeof(double));
  for (int j = 0; j < n; ++j) xa[j] = x[j] - alpha_bar * phi[j];
  orc_project(P, xa, px);
  for (int j = 0; j < n; ++j) xa[j] = x[j] - px[j];
  double r = dot(n, xa, phi) * (1.0 / alpha_bar);
  free(xa); free(px);
  return r > 0.0 ? r : 0.0;
}

/* BaseGeneralConProjector::stepLimit (MPRGPProjection.h:263-282). */
double orc_step_limit(const orc_projector *P, const double *x, const double *d, double alpha_cg) {
  double alpha = (alpha_cg == ORC_MAX) ? alpha_cg : (alpha_cg + ORC_EPS);
  double *jx = (double *)xcalloc((size_t)P->m, sizeof(double));
  double *jd = (double *)xcalloc((size_t)P->m, sizeof(double));
  j_mul(P, x, jx);
  j_mul(P, d, jd);
  for (int i = 0; i < P->m; ++i)
    if (jd[i] > ORC_EPS && jx[i] > P->c[i]) {
      double ti = (jx[i] - P->c[i]) / jd[i];
      alpha = (ti < alpha) ? ti : alpha; /* std::min<T>(alpha, ti) */
    }
  free(jx); free(jd);
  return alpha;
}

/* -------------------------------------------------------- preconditioner */

struct orc_precond {
  int n, m, enabled;
  double *invD;      /* 1/diag(A)                          */
  double *jminv;     /* J D^-1, same pattern as J          */
  double *inv_jmjt;  /* 1/diag(J D^-1 J^T)                 */
};

/* DiagonalInFacePreconSolver ctor (MPRGPPrecondition.h:42-51) and
 * DiagonalDecouplePreconSolver ctor (MPRGPPrecondition.h:78-90). */
orc_precond *orc_precond_new(const orc_csr *A, const orc_projector *P, int preconditioned) {
  orc_precond *pc = (orc_precond *)xcalloc(1, sizeof(orc_precond));
  pc->n = A->rows;
  pc->m = P->m;
  pc->enabled = preconditioned;
  pc->invD = (double *)xcalloc((size_t)pc->n, sizeof(double));
  if (preconditioned) {
    orc_get_diagonal(A, pc->invD);
    for (int i = 0; i < pc->n; ++i) pc->invD[i] = 1.0 / pc->invD[i];
  }
  int nnzj = P->m ? P->ptr[P->m] : 0;
  pc->jminv = (double *)xcalloc((size_t)nnzj, sizeof(double));
  pc->inv_jmjt = (double *)xcalloc((size_t)pc->m, sizeof(double));
  if (preconditioned && P->m > 0) {
    for (int i = 0; i < P->m; ++i) {
      double s = 0.0;
      for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) {
        pc->jminv[k] = P->val[k] * pc->invD[P->idx[k]];
        s += pc->jminv[k] * P->val[k];
      }
      pc->inv_jmjt[i] = 1.0 / s;
    }
  }
  return pc;
}

void orc_precond_free(orc_precond *pc) {
  if (!pc) return;
  free(pc->invD); free(pc->jminv); free(pc->inv_jmjt);
  free(pc);
}

/* DiagonalDecouplePreconSolver::solve (MPRGPPrecondition.h:91-105).  Note
 * that with no constraint rows (or preconditioning off) the result is phi,
 * i.e. the solve runs unpreconditioned. */
void orc_precond_solve(const orc_precond *pc, const orc_projector *P, const double *g,
                       double *z, const double *phi) {
  if (pc->enabled && pc->m > 0) {
    double *lam = P->lam;
    for (int i = 0; i < pc->m; ++i) {
      double s = 0.0;
      for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) s += pc->jminv[k] * g[P->idx[k]];
      lam[i] = pc->inv_jmjt[i] * s;
      if (P->face[i] == 0) lam[i] = 0.0;
    }
    double *t = P->tn2;
    memset(t, 0, sizeof(double) * (size_t)pc->n);
    for (int i = 0; i < pc->m; ++i)
      for (int k = P->ptr[i]; k < P->ptr[i + 1]; ++k) t[P->idx[k]] += pc->jminv[k] * lam[i];
    for (int j = 0; j < pc->n; ++j) z[j] = pc->invD[j] * g[j] - t[j];
  } else {
    if (z != phi) memcpy(z, phi, sizeof(double) * (size_t)pc->n);
  }
}

/* ---------------------------------------------------- power iteration */

static unsigned long long splitmix64(unsigned long long *s) {
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* MPRGPBase::specRad (MPRGPSolver.h:42-79).  ev is written only when the
 * iteration breaks early, exactly as in the reference (the previous iterate,
 * not the new one). */
double orc_spec_rad(const orc_csr *A, double *ev, int *ev_valid, double eps,
                    unsigned long long seed, int *iters_out) {
  int n = A->rows;
  double *tmp = (double *)xcalloc((size_t)n, sizeof(double));
  double *out = (double *)xcalloc((size_t)n, sizeof(double));
  if (ev && ev_valid && *ev_valid) {
    memcpy(tmp, ev, sizeof(double) * (size_t)n);
  } else {
    unsigned long long s = seed;
    for (int i = 0; i < n; ++i)
      tmp[i] = 2.0 * ((double)(splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    double nr = norm2(n, tmp);
    if (nr > 0.0)
      for (int i = 0; i < n; ++i) tmp[i] /= nr;
  }
  double nout = 1.0;
  int iter = 0;
  for (; iter <= 1000; iter++) {
    orc_spmv(A, tmp, out);
    nout = norm2(n, out);
    if (nout < ORC_EPS) {
      if (ev) { memcpy(ev, tmp, sizeof(double) * (size_t)n); if (ev_valid) *ev_valid = 1; }
      nout = ORC_EPS;
      break;
    }
    for (int i = 0; i < n; ++i) out[i] /= nout;
    double delta = 0.0;
    for (int i = 0; i < n; ++i) { double t = out[i] - tmp[i]; delta += t * t; }
    delta = sqrt(delta);
    if (delta <= eps) {
      if (ev) { memcpy(ev, tmp, sizeof(double) * (size_t)n); if (ev_valid) *ev_valid = 1; }
      break;
    }
    memcpy(tmp, out, sizeof(double) * (size_t)n);
  }
  if (iters_out) *iters_out = iter;
  free(tmp); free(out);
  return nout;
}

/* ------------------------------------------------------------ MPRGP core */

typedef struct {
  const orc_csr *A;
  const double *B;
  orc_projector *P;
  orc_precond *pc;
  int n;
  int max_iterations;
  double tol, gamma, alpha_bar;
  double *g, *p, *z, *beta, *phi, *gp, *y, *t1;
  int y_size;
  double residual_out;
  int result_code, num_cg, num_exp, numprop, iteration;
  orc_options *opt;  /* events and the f(P(y)) form (NULL: reference form, no events) */
  double *t2, *t3;   /* scratch of the closed-form objective */
} mprgp;

/* Appends one decision to the caller's event list (reads state only). */
static void record(mprgp *S, int kind, int outcome, double lhs, double rhs) {
  orc_options *o = S->opt;
  if (!o) return;
  if (o->events && o->n < o->cap) {
    orc_event *e = &o->events[o->n];
    e->iteration = S->iteration;
    e->kind = kind;
    e->outcome = outcome;
    e->pad = 0;
    e->lhs = lhs;
    e->rhs = rhs;
  }
  o->n++;
}

/* min over rows of (J y - c)_i, DBL_MAX without rows (for events only). */
static double min_slack(mprgp *S, const double *y) {
  double m = ORC_MAX;
  j_mul(S->P, y, S->P->lam);
  for (int i = 0; i < S->P->m; ++i) {
    double v = S->P->lam[i] - S->P->c[i];
    m = v < m ? v : m;
  }
  return m;
}

static void mprgp_init(mprgp *S) {
  int n = S->n;
  S->g = (double *)xcalloc((size_t)n, sizeof(double));
  S->p = (double *)xcalloc((size_t)n, sizeof(double));
  S->z = (double *)xcalloc((size_t)n, sizeof(double));
  S->beta = (double *)xcalloc((size_t)n, sizeof(double));
  S->phi = (double *)xcalloc((size_t)n, sizeof(double));
  S->gp = (double *)xcalloc((size_t)n, sizeof(double));
  S->y = (double *)xcalloc((size_t)n, sizeof(double));
  S->t1 = (double *)xcalloc((size_t)n, sizeof(double));
  S->t2 = (double *)xcalloc((size_t)n, sizeof(double));
  S->t3 = (double *)xcalloc((size_t)n, sizeof(double));
  S->y_size = 0;
}

static void mprgp_free(mprgp *S) {
  free(S->g); free(S->p); free(S->z); free(S->beta); free(S->phi); free(S->gp);
  free(S->y); free(S->t1); free(S->t2); free(S->t3);
}

/* g = A x - B */
static void grad(mprgp *S, const double *x) {
  orc_spmv(S->A, x, S->g);
  for (int i = 0; i < S->n; ++i) S->g[i] -= S->B[i];
}

/* MPRGPBase::initialize (MPRGPSolver.h:148-162). */
static void initialize(mprgp *S, const double *x) {
  grad(S, x);
  orc_decide_face(S->P, x);
  orc_phi(S->P, S->g, S->phi);
  orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
  S->result_code = -1;
  S->num_cg = S->num_exp = S->numprop = S->iteration = 0;
}

/* MPRGPBase::computeGradients (MPRGPSolver.h:163-173). */
static double compute_gradients(mprgp *S, int comp_phi) {
  if (comp_phi) orc_phi(S->P, S->g, S->phi);
  orc_beta(S->P, S->g, S->phi, S->beta);
  for (int i = 0; i < S->n; ++i) S->gp[i] = S->phi[i] + S->beta[i];
  return norm2(S->n, S->gp);
}

/* MPRGPBase::proportional (MPRGPSolver.h:174-181). */
static int proportional(mprgp *S, const double *x) {
  double left = dot(S->n, S->beta, S->beta);
  double phitphi = orc_phitphi(S->P, x, S->alpha_bar, S->phi);
  double right = S->gamma * S->gamma * phitphi;
  record(S, ORC_EV_PROP, left <= right, left, right);
  return left <= right;
}

/* MPRGPBase::CGStep (MPRGPSolver.h:182-195). */
static void cg_step(mprgp *S, const double *AP, double alpha_cg, double *x) {
  S->num_cg++;
  for (int i = 0; i < S->n; ++i) x[i] -= alpha_cg * S->p[i];
  for (int i = 0; i < S->n; ++i) S->g[i] -= alpha_cg * AP[i];
  orc_phi(S->P, S->g, S->phi);
  orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
  double b = dot(S->n, S->z, AP) / dot(S->n, S->p, AP);
  for (int i = 0; i < S->n; ++i) S->p[i] = S->z[i] - b * S->p[i];
}

/* MPRGPBase::ExpStep (MPRGPSolver.h:196-218); xTmp aliases the beta member. */
static void exp_step(mprgp *S, const double *AP, double alpha_f, double *x) {
  S->num_exp++;
  double *xt = S->beta;
  for (int i = 0; i < S->n; ++i) xt[i] = x[i] - alpha_f * S->p[i];
  for (int i = 0; i < S->n; ++i) S->g[i] -= alpha_f * AP[i];
  orc_decide_face(S->P, xt);
  orc_phi(S->P, S->g, S->phi);
  for (int i = 0; i < S->n; ++i) xt[i] -= S->alpha_bar * S->phi[i];
  orc_project(S->P, xt, x);
  grad(S, x);
  orc_decide_face(S->P, x);
  orc_phi(S->P, S->g, S->phi);
  orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
}

/* MPRGPBase::PropStep (MPRGPSolver.h:219-243); D aliases beta, AD aliases gp. */
static void prop_step(mprgp *S, double *x) {
  S->numprop++;
  double *D = S->beta, *AD = S->gp;
  orc_spmv(S->A, D, AD);
  double ddad = dot(S->n, D, AD);
  double alpha_cg = dot(S->n, S->g, D) / ddad;
  for (int i = 0; i < S->n; ++i) x[i] -= alpha_cg * D[i];
  double *xt = S->beta; /* xTmp = result (overwrites D, no longer needed) */
  memcpy(xt, x, sizeof(double) * (size_t)S->n);
  orc_project(S->P, xt, x);
  for (int i = 0; i < S->n; ++i) S->g[i] -= alpha_cg * AD[i];
  orc_decide_face(S->P, x);
  orc_phi(S->P, S->g, S->phi);
  orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
}

/* MPRGPBase::projectFunValue (MPRGPSolver.h:245-249). */
static double project_fun_value(mprgp *S, const double *x) {
  orc_project(S->P, x, S->t1);
  return orc_func_value(S->A, S->t1, S->B);
}

/* MPRGPBase::ExpMonotonicStep (MPRGPSolver.h:250-271). */
static void exp_monotonic_step(mprgp *S, double *x) {
  S->num_exp++;
  orc_project(S->P, x, S->y);
  S->y_size = S->n;
  grad(S, S->y);
  orc_decide_face(S->P, S->y);
  orc_phi(S->P, S->g, S->phi);
  for (int i = 0; i < S->n; ++i) S->y[i] -= S->alpha_bar * S->phi[i];
  orc_project(S->P, S->y, x);
  grad(S, x);
  orc_decide_face(S->P, x);
  orc_phi(S->P, S->g, S->phi);
  orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
  memcpy(S->p, S->z, sizeof(double) * (size_t)S->n);
}

/* f(P(y)) of the trial point y = x - alpha p from the step's sums (the
 * device mixed form's closed form; see ORC_F_CLOSED): f0 = f(x) at the
 * iterate, g = A x - B, AP = A p.  *fyu receives f(y) itself. */
static double closed_fy(mprgp *S, double f0, double alpha, double pd, const double *AP, double *fyu) {
  int n = S->n;
  orc_projector *P = S->P;
  *fyu = f0 - alpha * dot(n, S->p, S->g) + 0.5 * alpha * alpha * pd;
  /* d = P(y) - y on the rows violated by more than the feasibility tolerance */
  double *lam = P->lam, *d = S->t2, *ad = S->t3;
  j_mul(P, S->y, lam);
  for (int i = 0; i < P->m; ++i) {
    double v = P->c[i] - lam[i];
    lam[i] = (v > ORC_EPS) ? v / P->JJt[i] : 0.0;
  }
  jt_mul(P, lam, d);
  orc_spmv(S->A, d, ad);
  double share = 0.0;
  for (int i = 0; i < n; ++i) share += d[i] * ((S->g[i] - alpha * AP[i]) + 0.5 * ad[i]);
  return *fyu + share;
}

/* Closed forms tracked along the solve's own trajectory (traced solves):
 * f0 of the correct form continues from f(y) of each accepted step, f0 of
 * the planted wrong form from its own f(P(y)). */
typedef struct {
  double f0_fix, f0_bad;     /* f(x) of the next trial, per form      */
  double fp_fix, fyu_fix;    /* this trial: f(P(y)) and f(y), correct */
  double fp_bad;             /* this trial: f(P(y)), planted form      */
} closed_track;

/* f(P(y)) in the form the options ask for; with options, both closed forms
 * are evaluated too and recorded (ORC_EV_FY_CLOSED) next to the decision. */
static double trial_fy(mprgp *S, double alpha, double pd, const double *AP, closed_track *ct) {
  if (!S->opt) return project_fun_value(S, S->y);
  double fyu_bad;
  ct->fp_fix = closed_fy(S, ct->f0_fix, alpha, pd, AP, &ct->fyu_fix);
  ct->fp_bad = closed_fy(S, ct->f0_bad, alpha, pd, AP, &fyu_bad);
  switch (S->opt->fmode) {
    case ORC_F_CLOSED: return ct->fp_fix;
    case ORC_F_CLOSED_FROM_FPY: return ct->fp_bad;
    default: return project_fun_value(S, S->y);
  }
}

/* MPRGPBase::CGMonotonicStep (MPRGPSolver.h:272-316); AP aliases gp. */
static double cg_monotonic_step(mprgp *S, double *x, int *result_is_prop) {
  int n = S->n;
  double *AP = S->gp;
  orc_spmv(S->A, S->p, AP);
  double pd = dot(n, S->p, AP);
  double alpha_cg = dot(n, S->z, S->g) / pd;
  for (int i = 0; i < n; ++i) S->y[i] = x[i] - alpha_cg * S->p[i];
  S->y_size = n;
  double fx = orc_func_value(S->A, x, S->B);
  closed_track ct = {fx, fx, 0.0, 0.0, 0.0};
  double fy = trial_fy(S, alpha_cg, pd, AP, &ct);
  *result_is_prop = 1;
  for (;;) {
    if (!(*result_is_prop && S->residual_out > S->tol)) break;
    if (S->opt) {
      const double ms = min_slack(S, S->y);
      record(S, ORC_EV_CG_TEST, fy <= fx, fy, fx);
      record(S, ORC_EV_FY_CLOSED, 0, ct.fp_fix, ct.fp_bad);
      if (ms < -ORC_EPS) record(S, ORC_EV_Y_INFEASIBLE, 1, ms, -ORC_EPS);
    }
    if (!(fy <= fx && S->iteration < S->max_iterations)) break;
    record(S, ORC_EV_STEP, ORC_STEP_CG, 0.0, 0.0);
    /* the iterate becomes y, unprojected: the correct closed form continues
     * from f(y), the planted one from its f(P(y)) */
    ct.f0_fix = ct.fyu_fix;
    ct.f0_bad = ct.fp_bad;
    memcpy(x, S->y, sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) S->g[i] -= alpha_cg * AP[i];
    orc_phi(S->P, S->g, S->phi);
    orc_precond_solve(S->pc, S->P, S->g, S->z, S->phi);
    double b = dot(n, S->z, AP) / dot(n, S->p, AP);
    for (int i = 0; i < n; ++i) S->p[i] = S->z[i] - b * S->p[i];
    S->residual_out = compute_gradients(S, 0); /* overwrites gp (= AP) */
    record(S, ORC_EV_RESIDUAL, S->residual_out <= S->tol, S->residual_out, S->tol);
    *result_is_prop = proportional(S, x);
    orc_spmv(S->A, S->p, AP);
    pd = dot(n, S->p, AP);
    alpha_cg = dot(n, S->z, S->g) / pd;
    for (int i = 0; i < n; ++i) S->y[i] = x[i] - alpha_cg * S->p[i];
    fx = fy;
    fy = trial_fy(S, alpha_cg, pd, AP, &ct);
    S->num_cg++;
    S->iteration++;
  }
  return S->residual_out;
}

/* MPRGPMonotonic::solve (MPRGPSolver.h:412-454). */
static int solve_monotonic(mprgp *S, double *x) {
  initialize(S, x);
  for (S->iteration = 0; S->iteration < S->max_iterations;) {
    S->residual_out = compute_gradients(S, 1);
    record(S, ORC_EV_RESIDUAL, S->residual_out <= S->tol, S->residual_out, S->tol);
    if (S->residual_out <= S->tol) break;
    int result_is_prop = proportional(S, x);
    if (result_is_prop) S->residual_out = cg_monotonic_step(S, x, &result_is_prop);
    if (S->residual_out <= S->tol) break;
    int expand = result_is_prop || S->y_size <= 0;
    if (!expand) {
      const int feas = orc_is_feasible(S->P, S->y);
      if (S->opt) record(S, ORC_EV_FEASIBLE, feas, min_slack(S, S->y), -ORC_EPS);
      expand = !feas;
    }
    if (expand) {
      record(S, ORC_EV_STEP, ORC_STEP_EXP, 0.0, 0.0);
      exp_monotonic_step(S, x);
      S->iteration++;
    } else {
      record(S, ORC_EV_STEP, ORC_STEP_PROP, 0.0, 0.0);
      grad(S, x);
      prop_step(S, x);
      S->iteration++;
    }
  }
  memcpy(S->z, x, sizeof(double) * (size_t)S->n);
  orc_project(S->P, S->z, x);
  S->result_code = S->residual_out <= S->tol ? 0 : -1;
  return S->result_code;
}

/* MPRGP::solve (MPRGPSolver.h:356-396): classic variant with the
 * max-feasible-step test. */
static int solve_classic(mprgp *S, double *x) {
  initialize(S, x);
  for (S->iteration = 0; S->iteration < S->max_iterations; S->iteration++) {
    S->residual_out = compute_gradients(S, 1);
    if (S->residual_out <= S->tol) { S->result_code = 0; break; }
    if (proportional(S, x)) {
      double *AP = S->gp;
      orc_spmv(S->A, S->p, AP);
      double pd = dot(S->n, S->p, AP);
      double alpha_cg = dot(S->n, S->z, S->g) / pd;
      double alpha_f = orc_step_limit(S->P, x, S->p, alpha_cg);
      if (alpha_cg <= alpha_f) cg_step(S, AP, alpha_cg, x);
      else exp_step(S, AP, alpha_f, x);
    } else {
      prop_step(S, x);
    }
  }
  S->result_code = S->residual_out <= S->tol ? 0 : -1;
  return S->result_code;
}

/* MPRGPDecoupledCon::solve (MPRGPSolver.h:544-574) together with the
 * MPRGPBase constructor (MPRGPSolver.h:21-39). */
int orc_decoupled_solve(const orc_csr *A, const double *B, const orc_csr *J, const double *c,
                        double *x, double tol, int max_it, int variant, int preconditioned,
                        double alpha_bar, double *ev, int *ev_valid, int project_first,
                        orc_stats *st) {
  return orc_decoupled_solve_ex(A, B, J, c, x, tol, max_it, variant, preconditioned, alpha_bar, ev, ev_valid,
                                project_first, NULL, st);
}

int orc_decoupled_solve_ex(const orc_csr *A, const double *B, const orc_csr *J, const double *c,
                           double *x, double tol, int max_it, int variant, int preconditioned,
                           double alpha_bar, double *ev, int *ev_valid, int project_first,
                           orc_options *opt, orc_stats *st) {
  int n = A->rows;
  orc_projector *P = orc_projector_new(J, c);
  if (project_first) {
    double *x0 = (double *)xcalloc((size_t)n, sizeof(double));
    memcpy(x0, x, sizeof(double) * (size_t)n);
    orc_project(P, x0, x);
    free(x0);
  }
  orc_precond *pc = orc_precond_new(A, P, preconditioned);
  mprgp S;
  memset(&S, 0, sizeof(S));
  S.A = A; S.B = B; S.P = P; S.pc = pc; S.n = n;
  S.opt = opt;
  if (opt) opt->n = 0;
  S.max_iterations = max_it;
  S.tol = tol < 1e-30 ? 1e-30 : tol; /* setParameters (MPRGPSolver.h:34-39) */
  S.gamma = 1.0;
  int piters = 0;
  double lmax;
  if (alpha_bar > 0.0) {
    S.alpha_bar = alpha_bar;
    lmax = 2.0 / alpha_bar;
  } else {
    lmax = orc_spec_rad(A, ev, ev_valid, tol, 0x5EEDULL, &piters);
    S.alpha_bar = 2.0 / lmax;
  }
  S.residual_out = 0.0;
  mprgp_init(&S);
  int code = (variant == ORC_CLASSIC) ? solve_classic(&S, x) : solve_monotonic(&S, x);
  if (st) {
    st->code = code;
    st->iterations = S.iteration;
    st->num_cg = S.num_cg;
    st->num_exp = S.num_exp;
    st->num_prop = S.numprop;
    st->active = P->num_active;
    st->power_iters = piters;
    st->residual = S.residual_out;
    st->alpha_bar = S.alpha_bar;
    st->lambda_max = lmax;
  }
  mprgp_free(&S);
  orc_precond_free(pc);
  orc_projector_free(P);
  return code;
}
