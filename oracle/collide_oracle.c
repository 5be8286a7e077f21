/*
 * collide_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * Serial restatement of the reference's discrete plane collision in decoupled
 * mode: the rows J x >= c that one time step's collision handling hands to the
 * decoupled MPRGP solve.  Built with -ffp-contract=off, as the rest of the
 * oracle.
 *
 * Function -> reference (file:line) -> how it is pinned (tests/test_oracle_collide.py):
 *   orc_plane_collide  MprgpFemSolver.cpp:255-271 (reset, then every plane in order),
 *                      LinearConCollider.cpp:140-161 (handle(body, plane, delta)),
 *                      LinearConCollider.cpp:43-56 (addConPlane, 1e-4 duplicate test),
 *                      LinearConCollider.cpp:57-72, 206-230 (getConstraints: row = n on
 *                      the node's 3 columns, c = -p, rows in push order)
 *                      pinned by: brute-force first-hit rule per node, delta = +inf
 *                      (every node, first plane), exact ties n.x + p - delta = 0 (rows
 *                      kept, the test is <=), duplicate and near-duplicate planes
 *                      (no extra rows), the reference's normal assertion, and the
 *                      (plane, node) row order.
 */
#include <math.h>
#include <stdlib.h>

#include "mprgp_oracle.h"

/* One vertex's list of distinct planes seen this step (linear_con[vert_id]). */
typedef struct {
  int count;
  double *p; /* count x 4: nx, ny, nz, offset */
} plane_list;

/* addConPlane: index of a stored plane within 1e-4 (4-vector 2-norm), or -1
 * after appending the new plane. */
static int add_con_plane(plane_list *L, const double *pl) {
  for (int k = 0; k < L->count; ++k) {
    const double *q = L->p + 4 * k;
    double d0 = pl[0] - q[0], d1 = pl[1] - q[1], d2 = pl[2] - q[2], d3 = pl[3] - q[3];
    if (sqrt(d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3) <= 1e-4) return k;
  }
  double *q = L->p + 4 * L->count++;
  q[0] = pl[0];
  q[1] = pl[1];
  q[2] = pl[2];
  q[3] = pl[3];
  return -1;
}

int orc_plane_collide(int n_nodes, const double *x, int n_planes, const double *planes,
                      int *node, double *normal, double *c) {
  for (int k = 0; k < n_planes; ++k) {
    const double *pl = planes + 5 * k;
    double nn = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
    if (!(nn >= 1.0 - 1e-8 && nn <= 1.0 + 1e-8)) return -1; /* assert_in(n.norm(), ...) */
  }
  /* reset(num_verts): empty plane lists, no vertex collided */
  plane_list *lists = calloc(n_nodes ? (size_t)n_nodes : 1, sizeof(plane_list));
  double *store = calloc((size_t)(n_nodes ? n_nodes : 1) * (n_planes ? n_planes : 1) * 4, sizeof(double));
  char *collided = calloc(n_nodes ? (size_t)n_nodes : 1, 1);
  if (!lists || !store || !collided) abort();
  for (int i = 0; i < n_nodes; ++i) lists[i].p = store + (size_t)i * (n_planes ? n_planes : 1) * 4;
  int m = 0;
  for (int k = 0; k < n_planes; ++k) {
    const double *pl = planes + 5 * k;
    for (int i = 0; i < n_nodes; ++i) {
      const double *v = x + 3 * (size_t)i;
      /* n.dot(pos) + plane[3] - delta <= 0, Eigen's 3-term dot summed left to right */
      double d = pl[0] * v[0] + pl[1] * v[1];
      d = d + pl[2] * v[2];
      d = d + pl[3];
      d = d - pl[4];
      if (!(d <= 0.0)) continue;
      if (add_con_plane(&lists[i], pl) >= 0) continue;
      if (collided[i]) continue; /* decoupled: one row per vertex */
      collided[i] = 1;
      node[m] = i;
      normal[3 * m + 0] = pl[0];
      normal[3 * m + 1] = pl[1];
      normal[3 * m + 2] = pl[2];
      c[m] = -pl[3];
      ++m;
    }
  }
  free(lists);
  free(store);
  free(collided);
  return m;
}
