/*
 * ica_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Serial FP64 restatement of the reference's ICA comparator
 * (src/MPRGPSolver/ICASolver.h): ICASolver (:287-463, Otaduy et al.,
 * "Implicit Contact Handling for Deformable Objects", EG 2009) for
 *     A x = R + J^T lambda,  0 <= lambda  _|_  J x - p >= 0,
 * i.e. the KKT system of min 1/2 x'Ax - R'x s.t. J x >= p, with its inner
 * ProjectedGaussSeidel (GaussSeidel<true>, :20-150) on B = J Da^-1 J^T, as
 * ICAFemSolver::forward calls it (MprgpFemSolver.cpp:439-486).
 *
 * A = La + Da + Ua: Da the 3x3 diagonal blocks, La / Ua the entries outside
 * them below / above the diagonal (scalar index r > c / r < c; outside the
 * diagonal blocks that is node(r) > node(c) / <).  Each ICA iteration:
 *     c      = p + J Da^-1 (LaUa x - R)                       (:388)
 *     lambda = PGS(B, c) from the previous lambda               (:390)
 *     rhs    = J^T lambda - Ua x + R                            (:391-392)
 *     one forward block Gauss-Seidel sweep of (Da + La) x = rhs (:395-404)
 *     residual = |A x - R - J^T lambda|, stop when <= tol       (:406-411)
 *
 * Arithmetic order follows the reference's column-major Eigen products: a
 * product's entry for scalar row r adds the row's terms in ascending column
 * order (A and J in CSR with sorted columns give that order row by row), a
 * 3x3 block times a 3-vector adds its three terms in order, and J^T lambda
 * adds J's column entries in ascending row order.  Da^-1 is the adjugate
 * divided by the determinant (Eigen's 3x3 inverse is the same closed form;
 * its last bits may differ).
 *
 * Function -> reference -> pin (tests/test_oracle_ica.py):
 *   orc_pgs        GaussSeidel<true>::solve :68-111   brute-force minimiser of 1/2 l'Bl - c'l, l >= 0
 *   orc_ica_solve  ICASolver::solve :367-415          with no rows: A^-1 R (np.linalg.solve); one sweep
 *                                                     equals the block-lower-triangular solve (D+L) x1 =
 *                                                     R - U x0; with rows: the brute-force minimiser of
 *                                                     tiny coupled and decoupled QPs, lambda = their KKT
 *                                                     multipliers
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mprgp_oracle.h"

static void *icalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) abort();
  return p;
}

/* 3x3 inverse (row-major) by the adjugate: inv(i,j) = cof(j,i) / det, det
 * expanded along the first row. */
static void inv3(const double *m, double *inv) {
  const double c00 = m[4] * m[8] - m[5] * m[7];
  const double c01 = m[5] * m[6] - m[3] * m[8];
  const double c02 = m[3] * m[7] - m[4] * m[6];
  const double c10 = m[2] * m[7] - m[1] * m[8];
  const double c11 = m[0] * m[8] - m[2] * m[6];
  const double c12 = m[1] * m[6] - m[0] * m[7];
  const double c20 = m[1] * m[5] - m[2] * m[4];
  const double c21 = m[2] * m[3] - m[0] * m[5];
  const double c22 = m[0] * m[4] - m[1] * m[3];
  const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
  inv[0] = c00 / det; inv[1] = c10 / det; inv[2] = c20 / det;
  inv[3] = c01 / det; inv[4] = c11 / det; inv[5] = c21 / det;
  inv[6] = c02 / det; inv[7] = c12 / det; inv[8] = c22 / det;
}

void orc_inv3(const double *m, double *inv) { inv3(m, inv); }

/* GaussSeidel<true>::solve (:68-111): B = Lb + D + Ub by structural entries;
 * each sweep rhs = c - Ub lambda (old lambda), then row by row
 * rhs_r -= sum_{k<r} B_rk lambda_k (new lambda), lambda_r = max(0, rhs_r / B_rr);
 * residual |lambda - lambda_prev|, stop when <= tol.  B in CSR with sorted
 * columns.  Returns 1 when the tolerance was met, else 0; *iters receives the
 * reference's `iterations` (the sweep index at the stop, or max_it). */
int orc_pgs(const orc_csr *B, const double *c, double *lam, int max_it, double tol, int *iters, double *res) {
  const int m = B->rows;
  double *invd = (double *)icalloc((size_t)m, sizeof(double));
  double *rhs = (double *)icalloc((size_t)m, sizeof(double));
  double *pre = (double *)icalloc((size_t)m, sizeof(double));
  for (int r = 0; r < m; ++r)
    for (int k = B->ptr[r]; k < B->ptr[r + 1]; ++k)
      if (B->idx[k] == r) invd[r] = 1.0 / B->val[k];
  int it, succ = 0;
  double residual = -1.0;
  for (it = 0; it < max_it; ++it) {
    memcpy(pre, lam, sizeof(double) * (size_t)m);
    for (int r = 0; r < m; ++r) {
      double u = 0.0;
      for (int k = B->ptr[r]; k < B->ptr[r + 1]; ++k)
        if (B->idx[k] > r) u += B->val[k] * lam[B->idx[k]];
      rhs[r] = c[r] - u;
    }
    for (int r = 0; r < m; ++r) {
      for (int k = B->ptr[r]; k < B->ptr[r + 1]; ++k)
        if (B->idx[k] < r) rhs[r] -= lam[B->idx[k]] * B->val[k];
      const double v = rhs[r] * invd[r];
      lam[r] = v > 0 ? v : 0;
    }
    double s = 0.0;
    for (int r = 0; r < m; ++r) s += (lam[r] - pre[r]) * (lam[r] - pre[r]);
    residual = sqrt(s);
    if (residual <= tol) {
      succ = 1;
      break;
    }
  }
  if (iters) *iters = it;
  if (res) *res = residual;
  free(invd); free(rhs); free(pre);
  return succ;
}

/* B = J Da^-1 J^T (m x m, CSR, sorted columns, every structurally reachable
 * entry kept).  Entry (i, k) = sum over the nodes v shared by rows i and k,
 * in ascending v, of J_i(v)' Dinv_v J_k(v), each 3x3 product row by row. */
static void build_b(const orc_csr *J, int n_nodes, const double *dinv, int **bptr, int **bidx, double **bval) {
  const int m = J->rows;
  /* rows touching each node */
  int *cnt = (int *)icalloc((size_t)n_nodes + 1, sizeof(int));
  for (int i = 0; i < m; ++i) {
    int last = -1;
    for (int k = J->ptr[i]; k < J->ptr[i + 1]; ++k) {
      const int v = J->idx[k] / 3;
      if (v != last) cnt[v + 1]++;
      last = v;
    }
  }
  for (int v = 0; v < n_nodes; ++v) cnt[v + 1] += cnt[v];
  int *nrow = (int *)icalloc((size_t)cnt[n_nodes] + 1, sizeof(int));
  int *fill = (int *)icalloc((size_t)n_nodes, sizeof(int));
  for (int i = 0; i < m; ++i) {
    int last = -1;
    for (int k = J->ptr[i]; k < J->ptr[i + 1]; ++k) {
      const int v = J->idx[k] / 3;
      if (v != last) nrow[cnt[v] + fill[v]++] = i;
      last = v;
    }
  }
  /* dense 3-vector of row i on node v */
  double *jv = (double *)icalloc(3, sizeof(double));
  double *dense = (double *)icalloc((size_t)m, sizeof(double));
  char *mark = (char *)icalloc((size_t)m, 1);
  int *list = (int *)icalloc((size_t)m, sizeof(int));
  int cap = 16 * (m + 1), nnz = 0;
  int *ptr = (int *)icalloc((size_t)m + 1, sizeof(int));
  int *idx = (int *)icalloc((size_t)cap, sizeof(int));
  double *val = (double *)icalloc((size_t)cap, sizeof(double));
  for (int i = 0; i < m; ++i) {
    int nl = 0;
    /* the nodes of row i, ascending (J's columns are sorted) */
    for (int k = J->ptr[i]; k < J->ptr[i + 1];) {
      const int v = J->idx[k] / 3;
      double a[3] = {0.0, 0.0, 0.0};
      while (k < J->ptr[i + 1] && J->idx[k] / 3 == v) {
        a[J->idx[k] % 3] += J->val[k];
        ++k;
      }
      /* w = Dinv_v a */
      const double *D = dinv + 9 * (size_t)v;
      double w[3];
      for (int r = 0; r < 3; ++r) w[r] = D[3 * r] * a[0] + D[3 * r + 1] * a[1] + D[3 * r + 2] * a[2];
      for (int t = cnt[v]; t < cnt[v + 1]; ++t) {
        const int kr = nrow[t];
        jv[0] = jv[1] = jv[2] = 0.0;
        for (int q = J->ptr[kr]; q < J->ptr[kr + 1]; ++q)
          if (J->idx[q] / 3 == v) jv[J->idx[q] % 3] += J->val[q];
        const double d = jv[0] * w[0] + jv[1] * w[1] + jv[2] * w[2];
        if (!mark[kr]) {
          mark[kr] = 1;
          list[nl++] = kr;
          dense[kr] = d;
        } else {
          dense[kr] += d;
        }
      }
    }
    /* sort the row's columns (insertion: rows touch few others) */
    for (int a = 1; a < nl; ++a) {
      const int key = list[a];
      int b = a - 1;
      while (b >= 0 && list[b] > key) {
        list[b + 1] = list[b];
        --b;
      }
      list[b + 1] = key;
    }
    if (nnz + nl > cap) {
      while (nnz + nl > cap) cap *= 2;
      idx = (int *)realloc(idx, sizeof(int) * (size_t)cap);
      val = (double *)realloc(val, sizeof(double) * (size_t)cap);
      if (!idx || !val) abort();
    }
    for (int a = 0; a < nl; ++a) {
      idx[nnz] = list[a];
      val[nnz] = dense[list[a]];
      ++nnz;
      mark[list[a]] = 0;
    }
    ptr[i + 1] = nnz;
  }
  free(cnt); free(nrow); free(fill); free(jv); free(dense); free(mark); free(list);
  *bptr = ptr;
  *bidx = idx;
  *bval = val;
}

int orc_ica_b(const orc_csr *J, int n_nodes, const double *dinv, int *nnz_out, int **ptr, int **idx, double **val) {
  build_b(J, n_nodes, dinv, ptr, idx, val);
  *nnz_out = (*ptr)[J->rows];
  return 0;
}

void orc_ica_free(void *p) { free(p); }

/* ICASolver::reset(A, R) + solve(J, p, x) (:298-415).  A: n x n CSR, sorted
 * columns, n = 3 * nodes; J: m x n CSR, sorted columns.  x: in = the start,
 * out = the result; lambda (m, may be NULL) receives the multipliers.
 * Returns 0 when |A x - R - J^T lambda| <= tol was reached, else -1. */
int orc_ica_solve(const orc_csr *A, const double *R, const orc_csr *J, const double *p, double *x, double tol,
                  int max_it, int gs_max_it, double gs_tol, double *lambda, orc_ica_stats *st) {
  const int n = A->rows, nn = n / 3, m = J ? J->rows : 0;
  double *dinv = (double *)icalloc((size_t)9 * nn, sizeof(double));
  for (int v = 0; v < nn; ++v) {
    double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 3; ++a) {
      const int r = 3 * v + a;
      for (int k = A->ptr[r]; k < A->ptr[r + 1]; ++k)
        if (A->idx[k] / 3 == v) D[3 * a + A->idx[k] % 3] = A->val[k];
    }
    inv3(D, dinv + 9 * (size_t)v);
  }
  int *bptr = NULL, *bidx = NULL;
  double *bval = NULL;
  orc_csr B = {0, 0, NULL, NULL, NULL};
  if (m > 0) {
    build_b(J, nn, dinv, &bptr, &bidx, &bval);
    B.rows = B.cols = m;
    B.ptr = bptr;
    B.idx = bidx;
    B.val = bval;
  }
  double *lam = (double *)icalloc((size_t)m, sizeof(double));
  double *t = (double *)icalloc((size_t)n, sizeof(double));
  double *u = (double *)icalloc((size_t)n, sizeof(double));
  double *c = (double *)icalloc((size_t)m, sizeof(double));
  double *jtl = (double *)icalloc((size_t)n, sizeof(double));
  double *rhs = (double *)icalloc((size_t)n, sizeof(double));
  int it, succ = 0, gs_total = 0;
  double residual = -1.0, gs_res = -1.0;
  for (it = 0; it < max_it; ++it) {
    if (m > 0) {
      /* c = p + J Da^-1 (LaUa x - R) */
      for (int r = 0; r < n; ++r) {
        double s = 0.0;
        for (int k = A->ptr[r]; k < A->ptr[r + 1]; ++k)
          if (A->idx[k] / 3 != r / 3) s += A->val[k] * x[A->idx[k]];
        t[r] = s - R[r];
      }
      for (int v = 0; v < nn; ++v) {
        const double *D = dinv + 9 * (size_t)v;
        for (int a = 0; a < 3; ++a)
          u[3 * v + a] = D[3 * a] * t[3 * v] + D[3 * a + 1] * t[3 * v + 1] + D[3 * a + 2] * t[3 * v + 2];
      }
      for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int k = J->ptr[i]; k < J->ptr[i + 1]; ++k) s += J->val[k] * u[J->idx[k]];
        c[i] = p[i] + s;
      }
      int gi = 0;
      const int met = orc_pgs(&B, c, lam, gs_max_it, gs_tol, &gi, &gs_res);
      gs_total += gi + met;  /* sweeps run: a sweep that met the tolerance ends the loop uncounted */
    }
    /* J^T lambda, J's column entries in ascending row order */
    memset(jtl, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < m; ++i)
      for (int k = J->ptr[i]; k < J->ptr[i + 1]; ++k) jtl[J->idx[k]] += J->val[k] * lam[i];
    /* rhs = J^T lambda - Ua x + R (old x) */
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int k = A->ptr[r]; k < A->ptr[r + 1]; ++k)
        if (A->idx[k] / 3 > r / 3) s += A->val[k] * x[A->idx[k]];
      rhs[r] = (jtl[r] - s) + R[r];
    }
    /* the forward block sweep: lower entries with the new x */
    for (int v = 0; v < nn; ++v) {
      for (int a = 0; a < 3; ++a) {
        const int r = 3 * v + a;
        for (int k = A->ptr[r]; k < A->ptr[r + 1]; ++k)
          if (A->idx[k] / 3 < v) rhs[r] -= A->val[k] * x[A->idx[k]];
      }
      const double *D = dinv + 9 * (size_t)v;
      for (int a = 0; a < 3; ++a)
        x[3 * v + a] = D[3 * a] * rhs[3 * v] + D[3 * a + 1] * rhs[3 * v + 1] + D[3 * a + 2] * rhs[3 * v + 2];
    }
    /* residual |A x - R - J^T lambda| */
    double s2 = 0.0;
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int k = A->ptr[r]; k < A->ptr[r + 1]; ++k) s += A->val[k] * x[A->idx[k]];
      const double e = (s - R[r]) - jtl[r];
      s2 += e * e;
    }
    residual = sqrt(s2);
    if (residual <= tol) {
      succ = 1;
      break;
    }
  }
  if (lambda)
    for (int i = 0; i < m; ++i) lambda[i] = lam[i];
  if (st) {
    st->code = succ ? 0 : -1;
    st->iterations = it;
    st->gs_iterations = gs_total;
    st->residual = residual;
    st->gs_residual = gs_res;
  }
  free(dinv); free(bptr); free(bidx); free(bval);
  free(lam); free(t); free(u); free(c); free(jtl); free(rhs);
  return succ ? 0 : -1;
}
