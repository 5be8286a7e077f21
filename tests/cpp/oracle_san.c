/*
 * oracle_san.c -- runs every entry of the CPU oracle once on a small beam,
 * for a build under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_sanitizers.py).  Input file (text): N T, then N lines of rest
 * positions, N lines of current positions, N lines of velocities, T lines of
 * tets.  Prints one summary line; any memory or UB error aborts the process.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/mprgp_oracle.h"

static void *xmalloc(size_t n) {
  void *p = calloc(n ? n : 1, 1);
  if (!p) abort();
  return p;
}

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  FILE *f = fopen(argv[1], "r");
  if (!f) return 2;
  int N, T;
  if (fscanf(f, "%d %d", &N, &T) != 2 || N <= 0 || T <= 0) return 2;
  double *X0 = xmalloc(sizeof(double) * 3 * N), *xn = xmalloc(sizeof(double) * 3 * N),
         *vn = xmalloc(sizeof(double) * 3 * N);
  int *tets = xmalloc(sizeof(int) * 4 * T);
  for (int i = 0; i < 3 * N; ++i) if (fscanf(f, "%lf", &X0[i]) != 1) return 2;
  for (int i = 0; i < 3 * N; ++i) if (fscanf(f, "%lf", &xn[i]) != 1) return 2;
  for (int i = 0; i < 3 * N; ++i) if (fscanf(f, "%lf", &vn[i]) != 1) return 2;
  for (int i = 0; i < 4 * T; ++i) if (fscanf(f, "%d", &tets[i]) != 1) return 2;
  fclose(f);
  const int n = 3 * N;
  double *d = xmalloc(sizeof(double) * 9 * T), *vol = xmalloc(sizeof(double) * T),
         *mass = xmalloc(sizeof(double) * N), *young = xmalloc(sizeof(double) * T),
         *pois = xmalloc(sizeof(double) * T), *b = xmalloc(sizeof(double) * n);
  for (int t = 0; t < T; ++t) { young[t] = 1e5; pois[t] = 0.45; }
  orc_fem_setup(N, T, tets, X0, d, vol, mass);
  const double grav[3] = {0.0, -9.8, 0.0};
  int *ptr, *idx;
  double *val;
  if (orc_fem_build_system(N, T, tets, d, vol, mass, young, pois, 1000.0, xn, vn, grav, 0.01, ORC_LINEAR,
                           &ptr, &idx, &val, b))
    return 3;
  /* the corotated, damped form at a moved iterate exercises the polar step */
  double *x1 = xmalloc(sizeof(double) * n), *b2 = xmalloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) x1[i] = xn[i] + 1e-3 * vn[i];
  int *ptr2, *idx2;
  double *val2;
  if (orc_fem_build_system_damped(N, T, tets, d, vol, mass, young, pois, 1000.0, xn, vn, x1, grav, 0.01,
                                  ORC_COROTATIONAL, 0.01, 0.02, &ptr2, &idx2, &val2, b2))
    return 3;
  double *v = xmalloc(sizeof(double) * n);
  orc_implicit_euler_velocity(n, xn, vn, x1, 0.01, v);
  /* ground plane y >= 0 and a tilted wall, band 0.02 */
  const double planes[10] = {0.0, 1.0, 0.0, 0.0, 0.02, -0.9848077530122080, 0.1736481776669303, 0.0, 0.4, 0.02};
  int *node = xmalloc(sizeof(int) * N);
  double *nrm = xmalloc(sizeof(double) * 3 * N), *c = xmalloc(sizeof(double) * N);
  const int m = orc_plane_collide(N, xn, 2, planes, node, nrm, c);
  if (m < 0) return 4;
  /* J as CSR (one row per constrained node) */
  int *jp = xmalloc(sizeof(int) * (m + 1)), *jc = xmalloc(sizeof(int) * 3 * (m + 1));
  double *jv = xmalloc(sizeof(double) * 3 * (m + 1));
  for (int i = 0; i < m; ++i) {
    jp[i] = 3 * i;
    for (int k = 0; k < 3; ++k) { jc[3 * i + k] = 3 * node[i] + k; jv[3 * i + k] = nrm[3 * i + k]; }
  }
  jp[m] = 3 * m;
  orc_csr A = {n, n, ptr, idx, val}, J = {m, n, jp, jc, jv};
  double *ev = xmalloc(sizeof(double) * n);
  int ev_valid = 0, iters = 0;
  const double lam = orc_spec_rad(&A, ev, &ev_valid, 1e-6, 7, &iters);
  orc_stats st[3];
  int codes[3];
  for (int variant = 0; variant < 3; ++variant) {
    double *x = xmalloc(sizeof(double) * n);
    memcpy(x, xn, sizeof(double) * n);
    /* monotone with the power method from ev, classic, and the monotone
     * variant unpreconditioned */
    codes[variant] = orc_decoupled_solve(&A, b, &J, c, x, 1e-8, 2000, variant == 1 ? ORC_CLASSIC : ORC_MONOTONIC,
                                         variant != 2, variant == 0 ? 0.0 : 2.0 / lam, ev, &ev_valid, 1,
                                         &st[variant]);
    free(x);
  }
  /* the projector's and preconditioner's entries on their own */
  orc_projector *P = orc_projector_new(&J, c);
  double *y = xmalloc(sizeof(double) * n), *g = xmalloc(sizeof(double) * n), *phi = xmalloc(sizeof(double) * n),
         *be = xmalloc(sizeof(double) * n), *z = xmalloc(sizeof(double) * n);
  orc_project(P, xn, y);
  orc_decide_face(P, y);
  orc_spmv(&A, y, g);
  for (int i = 0; i < n; ++i) g[i] -= b[i];
  orc_phi(P, g, phi);
  orc_beta(P, g, phi, be);
  const double pp = orc_phitphi(P, y, 2.0 / lam, phi);
  const double sl = orc_step_limit(P, y, g, 1.0);
  orc_precond *pc = orc_precond_new(&A, P, 1);
  orc_precond_solve(pc, P, g, z, phi);
  const double fv = orc_func_value(&A, y, b);
  orc_get_diagonal(&A, z);
  /* the coupled-row solve and the ICA comparator on the same rows */
  orc_stats gst;
  double *xg = xmalloc(sizeof(double) * n), *lamv = xmalloc(sizeof(double) * (m + 1));
  memcpy(xg, xn, sizeof(double) * n);
  const int gcode = orc_general_solve(&A, b, &J, c, xg, 1e-8, 2000, 1, 2.0 / lam, NULL, NULL, 1, &gst);
  orc_ica_stats ist;
  memcpy(xg, xn, sizeof(double) * n);
  const int icode = orc_ica_solve(&A, b, &J, c, xg, 1e-8, 2000, 20, 1e-3, lamv, &ist);
  printf("oracle_san ok: N=%d T=%d rows=%d lambda=%.6e codes=%d,%d,%d iters=%d,%d,%d active=%d "
         "feasible=%d phitphi=%.3e step=%.3e f=%.6e general=%d/%d ica=%d/%d/%d\n",
         N, T, m, lam, codes[0], codes[1], codes[2], st[0].iterations, st[1].iterations, st[2].iterations,
         orc_num_active(P), orc_is_feasible(P, y), pp, sl, fv, gcode, gst.iterations, icode, ist.iterations,
         ist.gs_iterations);
  free(xg); free(lamv);
  orc_precond_free(pc);
  orc_projector_free(P);
  orc_free(ptr); orc_free(idx); orc_free(val); orc_free(ptr2); orc_free(idx2); orc_free(val2);
  free(X0); free(xn); free(vn); free(tets); free(d); free(vol); free(mass); free(young); free(pois); free(b);
  free(x1); free(b2); free(v); free(node); free(nrm); free(c); free(jp); free(jc); free(jv); free(ev);
  free(y); free(g); free(phi); free(be); free(z);
  return 0;
}
