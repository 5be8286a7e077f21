/*
 * mprgp_amd.h -- C ABI of the MI355X-native MPRGP contact-QP solver.
 *
 * Solves  min_x 1/2 x'Ax - b'x  s.t.  n_i . x_k(i) >= c_i  (one half-space row
 * per node at most), the decoupled contact QP of the reference simulator, on
 * one AMD Instinct MI355X (gfx950).  Plain pointers and sizes only; every
 * function returns an mprgp status code (0 = success) unless noted.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   MPRGPDecoupledCon<T>::solve          src/MPRGPSolver/MPRGPSolver.h:538-590
 *   MPRGPMonotonic / MPRGP solve loops   src/MPRGPSolver/MPRGPSolver.h:345-456
 *   FixedSparseMatrix                    src/MPRGPSolver/MPRGPSolver.h:459-500
 *   DecoupledConProjector                src/MPRGPSolver/MPRGPProjection.h:320-397
 *   DiagonalDecouplePreconSolver         src/MPRGPSolver/MPRGPPrecondition.h:72-110
 *   MPRGPBase::specRad (power_ev)        src/MPRGPSolver/MPRGPSolver.h:42-79
 *   MprgpFemSolver::buildLinearSystem    src/CollisionHandle/MprgpFemSolver.cpp:315-356
 *   FEMSystem::buildSystem               src/ElasticSolver/FEMSystem.cpp:167-192
 *   MaterialEnergy::eval (LINEAR/COROT)  src/ElasticSolver/FEMEnergy.cpp:263-304, 657-1057
 *   LinearConCollider::handle (planes)   src/CollisionHandle/LinearConCollider.cpp:140-161
 *   MPRGPGeneralCon / GeneralConProjector src/MPRGPSolver/MPRGPSolver.h:594-626,
 *                                        src/MPRGPSolver/MPRGPProjection.h:400-520
 *   ICASolver (the comparator)           src/MPRGPSolver/ICASolver.h:287-463
 *
 * Threading: a context owns one HIP stream; calls that use it are serialised
 * by the context (several threads may share a context).  An engine runs its
 * own worker thread and stream on a context; a QP handed to an engine belongs
 * to it until its ticket completes.  No host-global mutable state besides a
 * thread-local last-error string.
 */
#ifndef MPRGP_AMD_H
#define MPRGP_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference returns only 0 / -1) */
#define MPRGP_OK 0
#define MPRGP_NOT_CONVERGED (-1) /* reference's -1: ||phi+beta|| > tol after max_it */
#define MPRGP_EBADINPUT (-2)     /* invalid sizes/indices, coupled rows, zero normal   */
#define MPRGP_ENOMEM (-3)        /* device allocation failed                           */
#define MPRGP_EHIP (-4)          /* HIP runtime error                                  */
#define MPRGP_ENAN (-5)          /* non-finite residual reached                        */
#define MPRGP_ENODEV (-6)        /* no usable GPU                                      */
#define MPRGP_EBUDGET (-7)       /* HBM budget of the context exceeded                 */
#define MPRGP_ESTATE (-8)        /* object not ready (e.g. rhs not set) or busy        */
#define MPRGP_EPENDING (-9)      /* engine ticket still running (wait timed out)       */

/* solver variants */
#define MPRGP_MONOTONIC 0 /* MPRGPMonotonic (MPRGPSolver.h:400-456), the simulator's */
#define MPRGP_CLASSIC 1   /* MPRGP with max-feasible-step test (MPRGPSolver.h:344-398) */

/* FEM materials (MaterialEnergy TYPE values used on the path) */
#define MPRGP_LINEAR 1
#define MPRGP_COROTATIONAL 2
/* flag or-ed into the material of mprgp_fem_assemble: write b only and keep
 * the matrix of the last full assembly (lagged stiffness between
 * re-assemblies).  COROTATIONAL forces then use the element rotations of that
 * assembly, so b and the kept matrix are one linearisation. */
#define MPRGP_RHS_ONLY 0x100

/* precision forms */
#define MPRGP_FP64 0  /* every vector and reduction in FP64                        */
#define MPRGP_MIXED 1 /* FP32 SpMV/updates, FP64 step lengths, KKT norm, refinement */

typedef struct mprgp_ctx mprgp_ctx;
typedef struct mprgp_matrix mprgp_matrix;
typedef struct mprgp_qp mprgp_qp;
typedef struct mprgp_fem mprgp_fem;

typedef struct {
  double tol;          /* bound on ||phi + beta||_2, absolute (reference tol)        */
  int max_it;          /* reference max_it                                           */
  int variant;         /* MPRGP_MONOTONIC (default) or MPRGP_CLASSIC                 */
  int precision;       /* MPRGP_FP64 (default) or MPRGP_MIXED                        */
  int preconditioned;  /* 1: tangent-Jacobi (reference `true`), 0: none              */
  int project_first;   /* 1: x = P(x) before solving (the J,c overload, :561-574)    */
  int use_ev;          /* 1: warm-start the power method from the QP's stored vector */
  double alpha_bar;    /* > 0: used as is; <= 0: 2/specRad(A, ev, power_tol)         */
  double power_tol;    /* power-method tolerance; <= 0 means tol, as the reference   */
  int tiles;           /* SELL slices per workgroup (1,2,4,8); 0 = autotuned         */
  int graph_steps;     /* launches per captured hipGraph; 0 = autotuned              */
  int max_refine;      /* MIXED: outer refinement sweeps (0 = default 8)             */
  int tail_tiles;      /* launches of at most this many tiles run the step kernel's  */
                       /* tail form (same results); 0 = 256, one per CU              */
  int xcd_layout;      /* workgroups over the 8 XCDs (same results): 0 runs of 4     */
                       /* tiles rotating, 1 one contiguous chunk of each QP per XCD  */
  int slices_per_wave; /* SELL slices (64 rows) each wavefront streams per launch:   */
                       /* 0 or 1, or 2 (rows per wavefront = 64 x this; a tile is    */
                       /* tiles x this slices; like tiles, it groups the sums)       */
  int from_x0;         /* 1: start from the QP's stored warm start (its last set_x), */
                       /* as mprgp_qp_reset_x would; a batch restarts all such QPs */
                       /* in one launch                                            */
  int reserved_i[1];   /* set to 0                                                   */
  double inner_tol;    /* MIXED: inner tolerance factor (0 = default 1e-3 relative)  */
  double reserved_d[3];
} mprgp_params;

typedef struct {
  int code;           /* 0 converged, -1 not converged (reference result_code)      */
  int iterations;     /* reference `iteration` ("total iter")                       */
  int num_cg, num_exp, num_prop;
  int active;         /* COUNT_CONSTRAINTS(face) at exit                             */
  int power_iters;    /* power-method iterations run in this solve                  */
  int ev_valid;       /* power vector stored for the next solve                     */
  int refinements;    /* MIXED: outer refinement sweeps                             */
  int steps;          /* device launches this QP took part in                       */
  int launches;       /* step-kernel launches issued for the whole call             */
  int fp64_finish;    /* MIXED: 1 when every refinement sweep ran and the residual  */
                      /* was still above tol (the FP32 inner solve's floor), so the */
                      /* FP64 form finished the solve from that x; then iterations  */
                      /* and the step counts cover both                             */
  double residual;    /* final ||phi + beta||_2 (FP64)                              */
  double alpha_bar;
  double lambda_max;  /* spectral-radius estimate behind alpha_bar                  */
  double elapsed_ms;  /* device time of the solve call (whole batch)                */
  double queue_ms;    /* engine: wait from submission to admission                  */
  /* No counterpart in the reference: a bound on |x - x*|_2, the distance of the
   * returned x from the QP's exact minimiser, residual / lambda_min_lower with
   * lambda_min_lower a lower bound on A's smallest eigenvalue
   * (mprgp_matrix_set_lambda_min; FEM assembly sets it); -1 when none is known.
   * For x feasible with its face exactly the rows it meets, strong convexity
   * and the rows' nonnegative multipliers give
   * lambda_min |x - x*|^2 <= g(x)'(x - x*) <= |phi + beta| |x - x*|.  It is
   * the mixed form's per-solve error report: residual is its FP64 KKT norm. */
  double error_bound;
  double reserved_d[2];
} mprgp_stats;

/* ---- context ------------------------------------------------------------ */
void mprgp_params_default(mprgp_params *p);
const char *mprgp_last_error(void);
int mprgp_device_count(int *count);
/* hbm_budget_bytes = 0: no budget beyond the device's memory. */
int mprgp_ctx_create(int device, size_t hbm_budget_bytes, mprgp_ctx **out);
void mprgp_ctx_destroy(mprgp_ctx *ctx);
int mprgp_ctx_memory(mprgp_ctx *ctx, size_t *in_use, size_t *budget);

/* ---- system matrix: FixedSparseMatrix (MPRGPSolver.h:459-500) ------------- */
/* 3x3 block CSR in node order: rowptr[n_nodes+1], colidx[nb], blocks[9*nb]
 * row-major.  Every row needs its diagonal block.  The matrix is copied into
 * device memory (SELL-64 layout in a row-length order) and owned by ctx. */
int mprgp_matrix_create_bsr3(mprgp_ctx *ctx, int n_nodes, const int *rowptr, const int *colidx,
                             const double *blocks, mprgp_matrix **out);
/* Scalar CSR of a symmetric matrix with n = 3*nodes rows.  For a symmetric
 * Eigen::SparseMatrix<double> (column-major) pass outerIndexPtr(),
 * innerIndexPtr(), valuePtr(): its CSC arrays are the CSR arrays. */
int mprgp_matrix_create_csr(mprgp_ctx *ctx, int n, const int *ptr, const int *idx,
                            const double *val, mprgp_matrix **out);
/* The matrix back as the node-order block CSR it was created from (same
 * block order; values as last written, e.g. by mprgp_fem assembly):
 * rowptr[n_nodes+1], colidx[nb], blocks[9*nb] row-major (sizes from
 * mprgp_matrix_info).  What the simulator's per-frame QP dump needs
 * (saveQP, MprgpFemSolver.h:129-133, with mprgp_write_qp). */
int mprgp_matrix_get_bsr3(mprgp_matrix *A, int *rowptr, int *colidx, double *blocks);
/* new values, same pattern and block order as at creation (for a matrix made
 * from scalar CSR: the block order mprgp_matrix_get_bsr3 returns).  A
 * non-finite value gives MPRGP_EBADINPUT, naming the first, and leaves the
 * values unchanged. */
int mprgp_matrix_update_bsr3(mprgp_matrix *A, const double *blocks);
int mprgp_matrix_info(const mprgp_matrix *A, int *n_nodes, long long *n_blocks, long long *bytes);
/* y = A x on host vectors (node order), FixedSparseMatrix::multiply (:469-473) */
int mprgp_matrix_spmv(mprgp_matrix *A, const double *x, double *y);
/* 1/2 x'Ax - x'b, FixedSparseMatrix::funcValue (:475-482), host vectors in node order */
int mprgp_matrix_func_value(mprgp_matrix *A, const double *x, const double *b, double *f);
/* No counterpart in the reference: declares lower > 0 a lower bound on A's
 * smallest eigenvalue (for A = M + h^2 K with K positive semidefinite, the
 * smallest lumped mass), which solves turn into mprgp_stats.error_bound;
 * 0 withdraws it.  mprgp_fem assembly sets it to min(m_i) (1 + h alpha_k);
 * new values (mprgp_matrix_update_bsr3) withdraw it. */
int mprgp_matrix_set_lambda_min(mprgp_matrix *A, double lower);
/* diag(A), 3 per node in node order: getDiagonal (MPRGPUtility.h:503-515), FixedSparseMatrix::diag */
int mprgp_matrix_diagonal(mprgp_matrix *A, double *diag);
void mprgp_matrix_destroy(mprgp_matrix *A);

/* ---- QP: matrix + rows + right-hand side + iterate, device-resident ------- */
int mprgp_qp_create(mprgp_matrix *A, mprgp_qp **out);
void mprgp_qp_destroy(mprgp_qp *qp);
/* DecoupledConProjector(J, JJt, c) (MPRGPProjection.h:330-331) in per-node
 * form: row i constrains node[i] with normal[3i..3i+2] and offset c[i].
 * Duplicate nodes (a second row on a node, even with disjoint DoFs as in
 * J = I), zero normals or bad indices give MPRGP_EBADINPUT. */
int mprgp_qp_set_rows(mprgp_qp *qp, int m, const int *node, const double *normal, const double *c);
/* Same from the sparse J the reference builds (LinearConCollider.cpp:206-230):
 * J is m x 3*nodes in CSR (col_major = 0) or CSC (col_major = 1, Eigen's
 * default).  Each row must lie inside one node's three columns. */
int mprgp_qp_set_rows_sparse(mprgp_qp *qp, int m, int col_major, const int *outer, const int *inner,
                             const double *val, const double *c);
/* One time step's discrete plane collision, decoupled (MprgpFemSolver.cpp:255-271
 * calling LinearConCollider::handle(body, plane, delta), LinearConCollider.cpp:140-161,
 * then getConstraints, :206-230), on the device from the QP's current x (the
 * last set_x or solve): every node with n.x + p - delta <= 0 for some plane
 * gets one row, that of the first such plane in the given order, J row = n
 * on the node, c = -p.  Replaces the QP's rows.  planes: n_planes x 5 doubles
 * {nx, ny, nz, p, delta}; delta may be +inf.  A normal whose length is not 1
 * within 1e-8 (the reference's assertion, :145) gives MPRGP_EBADINPUT; no x
 * set gives MPRGP_ESTATE.  *m_out (may be NULL) receives the row count. */
int mprgp_qp_collide_planes(mprgp_qp *qp, int n_planes, const double *planes, int *m_out);
/* The projector and preconditioner of the QP's rows, applied once to host
 * vectors in node order with the step kernel's own per-node code; the face
 * (active set) is the QP's, as the projector keeps it: set by decide_face or
 * by the last solve, all free for a new QP.
 *   project      DecoupledConProjector::project   MPRGPProjection.h:333-348
 *   is_feasible  BaseGeneralConProjector::isFeasible  :298-310 (1 or 0)
 *   decide_face  DECIDE_FACE :283-297 (*active may be NULL; COUNT_CONSTRAINTS)
 *   phi, beta    PHI :349-367, BETA :368-383 on the QP's face
 *   phitphi      PHITPHI :384-393 (alpha_bar > 1e-9, as it asserts)
 *   step_limit   stepLimit :263-282 (alpha_cg = DBL_MAX means none)
 *   precond      DiagonalDecouplePreconSolver::solve, MPRGPPrecondition.h:91-105,
 *                with 1/diag(A) of the QP's matrix; z = phi when the QP has no rows */
int mprgp_qp_project(mprgp_qp *qp, const double *x, double *y);
int mprgp_qp_is_feasible(mprgp_qp *qp, const double *x, int *feasible);
int mprgp_qp_decide_face(mprgp_qp *qp, const double *x, int *active);
int mprgp_qp_phi(mprgp_qp *qp, const double *g, double *phi);
int mprgp_qp_beta(mprgp_qp *qp, const double *g, const double *phi, double *beta);
int mprgp_qp_phitphi(mprgp_qp *qp, const double *x, double alpha_bar, const double *phi, double *out);
int mprgp_qp_step_limit(mprgp_qp *qp, const double *x, const double *d, double alpha_cg, double *alpha_f);
int mprgp_qp_precond(mprgp_qp *qp, const double *g, const double *phi, double *z);
/* The QP's rows in node order (ascending user node; the reference lists them
 * plane by plane): *m receives the count; node/normal/c hold cap rows (normal
 * 3 per row); with all three NULL only *m is written.  More rows than cap
 * gives MPRGP_EBADINPUT with *m set. */
int mprgp_qp_get_rows(mprgp_qp *qp, int cap, int *m, int *node, double *normal, double *c);
int mprgp_qp_set_rhs(mprgp_qp *qp, const double *b);
int mprgp_qp_get_rhs(mprgp_qp *qp, double *b);       /* node order               */
int mprgp_qp_set_x(mprgp_qp *qp, const double *x);   /* also kept as the reset point */
int mprgp_qp_reset_x(mprgp_qp *qp);                  /* x = last set_x, on device   */
int mprgp_qp_get_x(mprgp_qp *qp, double *x);
int mprgp_qp_set_ev(mprgp_qp *qp, const double *ev); /* power_ev warm start (:47-48)*/
int mprgp_qp_get_ev(mprgp_qp *qp, double *ev, int *valid);
int mprgp_qp_get_face(mprgp_qp *qp, unsigned char *face_per_node);
/* device launches the last solve spent in each internal operation (counts[0..n)) */
int mprgp_qp_op_counts(mprgp_qp *qp, int *counts, int n);
/* MPRGPDecoupledCon<double>::solve<FixedSparseMatrix, preconditioned>
 * (MPRGPSolver.h:544-559): x is read and overwritten on the device. */
int mprgp_qp_solve(mprgp_qp *qp, const mprgp_params *params, mprgp_stats *stats);
/* No counterpart in the reference: mprgp_qp_solve for callers that cannot
 * allocate (serving loops that must not call hipMalloc, harnesses that forbid
 * it).  The same solve, bit for bit, with no device or host allocation: its
 * launch tables live in a static per-GPU scratch array of the library
 * (657 KB, held until the process ends) and the step kernel is launched
 * directly, without graph capture.  Limits: QPs of at most 16384 workgroup
 * slots (2^20 nodes at params->tiles = 1, 2^22 at the default shape), else
 * MPRGP_EBADINPUT; the mixed form needs the QP's FP32 storage from an earlier
 * mprgp_qp_solve (else MPRGP_ESTATE).  Calls on one GPU share the array and
 * are serialised within the process; concurrent calls from several processes
 * each have their own copy. */
int mprgp_qp_solve_noalloc(mprgp_qp *qp, const mprgp_params *params, mprgp_stats *stats);
/* Many QPs advanced together (params: one per QP, or NULL-terminated use of
 * params[0] for all when params_stride == 0).  Returns MPRGP_OK when every QP
 * converged, MPRGP_NOT_CONVERGED if any did not, an error otherwise. */
int mprgp_solve_batch(mprgp_ctx *ctx, int count, mprgp_qp *const *qps, const mprgp_params *params,
                      int params_stride, mprgp_stats *stats);

/* ---- coupled rows -----------------------------------------------------------
 * MPRGPGeneralCon<double>::solve<FixedSparseMatrix, preconditioned>(A, B, J,
 * c, x, tol, max_it) (MPRGPSolver.h:611-623) with GeneralConProjector
 * (MPRGPProjection.h:400-520): rows J x >= c that may span several nodes,
 * such as the self-contact rows of LinearConCollider.cpp:186-217 (a vertex
 * against a triangle, 12 nonzeros); J is m x 3*nodes CSR in user DoF order,
 * m <= 2^20, at most 2048 rows in one connected component of J J^T's
 * pattern (rows linked through shared nodes), with J J^T nonsingular on
 * every set of rows a projection frees.  Solves the QP's matrix, right-hand side and x (projected first when
 * params->project_first; the QP's own per-node rows are not used); FP64,
 * monotone variant; params->preconditioned selects the reference's
 * DiagonalInFacePreconSolver (z = D^-1 phi).  alpha_bar <= 0: the power
 * method from the QP's stored ev (params->use_ev) or the start vector.
 * Returns 0, MPRGP_NOT_CONVERGED, or MPRGP_EBADINPUT when a projection's QP
 * or the active rows' factorisation fails.  stats->error_bound is -1. */
int mprgp_general_solve(mprgp_qp *qp, int m, const int *j_ptr, const int *j_col, const double *j_val,
                        const double *c, const mprgp_params *params, mprgp_stats *stats);

/* ---- the ICA comparator ------------------------------------------------------
 * ICASolver(max_it, tol, gs_max_it, gs_tol) .reset(A, b) .solve(J, p, x)
 * (ICASolver.h:287-415), as ICAFemSolver::forward runs it
 * (MprgpFemSolver.cpp:439-486): the solver the reference compares MPRGP
 * with, for A x = b + J^T lambda, 0 <= lambda _|_ J x - p >= 0, on the QP's
 * matrix, right-hand side and x (the start, overwritten; the reference
 * starts from the unconstrained solution).  Rows J x >= p as for
 * mprgp_general_solve (CSR over the user's DoFs, any coupling; the QP's own
 * per-node rows are not used).  Each iteration: c = p + J Da^-1 (LaUa x - b),
 * lambda by projected Gauss-Seidel on J Da^-1 J^T (at most gs_max_it sweeps,
 * to gs_tol, from the previous lambda), then one forward 3x3-block
 * Gauss-Seidel sweep in the user's node order; it stops when
 * |A x - b - J^T lambda| <= tol.  lambda (m, may be NULL) receives the
 * multipliers and *pgs_sweeps (may be NULL) the projected Gauss-Seidel
 * sweeps run.  stats: code (0 met, -1 not), iterations (the reference's
 * counter), residual, elapsed_ms; error_bound -1.  Returns MPRGP_OK or
 * MPRGP_NOT_CONVERGED. */
int mprgp_ica_solve(mprgp_qp *qp, int m, const int *j_ptr, const int *j_col, const double *j_val,
                    const double *p, double tol, int max_it, int gs_max_it, double gs_tol, double *lambda,
                    int *pgs_sweeps, mprgp_stats *stats);

/* ---- one-shot drop-in ------------------------------------------------------
 * MPRGPDecoupledCon<double>::solve<FixedSparseMatrix<double>,true>(A, B, J, c,
 * x, tol, max_it, name, ev) (MPRGPSolver.h:561-574): A as symmetric scalar
 * CSR/CSC arrays (n x n, n = 3*nodes), J in CSC (Eigen column-major, m x n),
 * x projected first, ev/ev_valid the caller's power vector (may be NULL). */
int mprgp_decoupled_solve(int device, int n, const int *a_outer, const int *a_inner,
                          const double *a_val, const double *B, int m, const int *j_outer,
                          const int *j_inner, const double *j_val, const double *c, double *x,
                          double tol, int max_it, double *ev, int *ev_valid, mprgp_stats *stats);

/* ---- QP files ----------------------------------------------------------------
 * The reference's binary QP file of A x = B s.t. J x >= c with initial value
 * x0 (writeQP / loadQP, MPRGPUtility.h:406-487): A and J as {size_t rows,
 * size_t cols, size_t nnz, nnz x Eigen::Triplet<double> (int32 row, int32
 * col, double value)}, then B, c and x0 as doubles, in the order A, B, J, c,
 * x0.  The readers return CSR arrays with columns ascending and duplicate
 * triplets summed (setFromTriplets); a file that cannot be read or is
 * malformed gives MPRGP_EBADINPUT (the reference's solve returns -1). */
/* n = A's rows, m = J's rows, the CSR nonzero counts of A and J */
int mprgp_qp_file_info(const char *file_name, int *n, int *m, long long *a_nnz, long long *j_nnz);
/* a_ptr[n+1], a_col/a_val[a_nnz], B[n], j_ptr[m+1], j_col/j_val[j_nnz], c[m], x0[n] */
int mprgp_qp_file_read(const char *file_name, int *a_ptr, int *a_col, double *a_val, double *B,
                       int *j_ptr, int *j_col, double *j_val, double *c, double *x0);
/* writeQP (MPRGPUtility.h:406-443) from CSR A (n x n) and CSR J (m x n) */
int mprgp_write_qp(const char *file_name, int n, const int *a_ptr, const int *a_col, const double *a_val,
                   const double *B, int m, const int *j_ptr, const int *j_col, const double *j_val,
                   const double *c, const double *x0);
/* MPRGPDecoupledCon<double>::solve(file_name, x, tol, max_it, name, ev)
 * (MPRGPSolver.h:576-589): reads the file, then solves as
 * mprgp_decoupled_solve from its x0 (projected first).  x (capacity n_x)
 * receives the solution; *n_out (may be NULL) the QP's size n. */
int mprgp_decoupled_solve_file(int device, const char *file_name, double *x, int n_x, int *n_out,
                               double tol, int max_it, double *ev, int *ev_valid, mprgp_stats *stats);

/* ---- FEM assembly on the device ------------------------------------------
 * FEMSystem::buildSystem (FEMSystem.cpp:167-192) + MprgpFemSolver::
 * buildLinearSystem (MprgpFemSolver.cpp:315-356) for one body at the first
 * Newton iterate x1 = x_n:  A = M + h^2 K(x_n),
 * b = h^2 f(x_n) + h^2 M g + M (x_n + h v_n - x_n) + A x_n,
 * K from LINEAR or COROTATIONAL tets (MaterialEnergy::eval, FEMEnergy.cpp:263-304,
 * 657-1057), lumped mass rho*sum(V/4) (FEMMesh.cpp:488-510).  The mesh is
 * fixed at creation; tets are re-oriented to positive volume (makePositive,
 * FEMMesh.cpp:158-165).  young/poisson per tet. */
int mprgp_fem_create(mprgp_ctx *ctx, int n_nodes, int n_tets, const int *tets, const double *rest_pos,
                     const double *young, const double *poisson, double density, mprgp_fem **out);
/* Writes A into the FEM's matrix and b into qp (created on mprgp_fem_matrix).
 * x_n, v_n: node positions and velocities (node order, host); gravity[3]. */
int mprgp_fem_assemble(mprgp_fem *fem, int material, const double *x_n, const double *v_n,
                       const double *gravity, double h, mprgp_qp *qp);
mprgp_matrix *mprgp_fem_matrix(mprgp_fem *fem);
void mprgp_fem_destroy(mprgp_fem *fem);
/* Rayleigh damping of the body (FEMSystem::setDamping, FEMSystem.cpp:122-126,
 * as Simulator.cpp:159-167 sets it from the scene's alpha_k and alpha_m; a
 * new FEM has none, as a new FEMSystem).  Assemblies then write
 * A = M (1 + h alpha_k) + K (h^2 + h alpha_m) and add
 * alpha_m K (x_n - x1) h + M (x_n - x1) h alpha_k to b (buildSystem,
 * FEMSystem.cpp:167-192, with CRHS = (x_n - x1) h). */
int mprgp_fem_set_damping(mprgp_fem *fem, double alpha_k, double alpha_m);
/* mprgp_fem_assemble at a Newton iterate x1 (buildLinearSystem after
 * updatePos, MprgpFemSolver.cpp:315-356): elements evaluated at x1,
 * b = h^2 f(x1) + h^2 M g + M (x_n + h v_n - x1) + A x1.  x1 = NULL means x_n. */
int mprgp_fem_assemble_at(mprgp_fem *fem, int material, const double *x_n, const double *v_n,
                          const double *x1, const double *gravity, double h, mprgp_qp *qp);

/* ---- time step: MprgpFemSolver::advance (MprgpFemSolver.cpp:193-203) -----
 * One body's positions and velocities kept on the device (node order), with
 * the simulator's integrator, implicit Euler (Simulator.cpp:313 calling
 * FEMSolver::resetImplicitEuler, FEMSystem.h:140-142: beta = 1, beta2 = 0.5,
 * gamma = 1), no damping.  A step is the planes' collision at x_n
 * (handleCollDetection/DCDhandle, :255-271), DecoupledMprgpFemSolver::forward's
 * Newton loop (:488-538: assemble at x1, x = P(x1), MPRGP solve, updatePos
 * :358-375, stop when max|x - x1| < eps or after maxIter solves), then
 * updateMesh (:377-390): a = (x - (x_n + h v_n)) / (h h), v = v_n + a h,
 * x_n = x. */
typedef struct {
  double h;               /* time step                                              */
  double gravity[3];
  int material;           /* MPRGP_LINEAR or MPRGP_COROTATIONAL, | MPRGP_RHS_ONLY  */
  int n_planes;           /* planes: n_planes x {nx, ny, nz, p, delta}, unit normals */
  const double *planes;
  int newton_max_it;      /* forward's "maxIter" (resetImplicitEuler default 10)    */
  double newton_eps;      /* forward's "eps" on max|x - x1| (default 1e-4)          */
  int reserved_i[4];      /* set to 0                                               */
  double reserved_d[4];
} mprgp_step_params;

typedef struct {
  int newton_iters;       /* solves run                                             */
  int rows;               /* constraint rows of the step's collision                */
  int code;               /* MPRGP_OK, or MPRGP_NOT_CONVERGED if any solve was not  */
  int iterations;         /* MPRGP iterations over the step's solves                */
  double delta;           /* last max|x - x1| (updatePos's tol_err)                 */
  mprgp_stats last;       /* the last solve's statistics                            */
} mprgp_step_stats;

int mprgp_fem_set_state(mprgp_fem *fem, const double *x, const double *v);
int mprgp_fem_get_state(mprgp_fem *fem, double *x, double *v); /* either may be NULL */
/* The whole step, solving with mprgp_qp_solve (params->project_first is
 * forced to 1, as forward projects x1 before each solve).  Returns
 * stats->code; the state advances also when a solve did not converge. */
int mprgp_fem_advance(mprgp_fem *fem, mprgp_qp *qp, const mprgp_step_params *sp,
                      const mprgp_params *params, mprgp_step_stats *stats);
/* A one-solve step split around a solve made elsewhere (an engine):
 * begin_step collides the planes at x_n (*rows, may be NULL), assembles at
 * x1 = x_n and sets the QP's x to x_n; end_step takes the QP's x as the new
 * positions, returns max|x - x_n| in *delta (may be NULL) and updates the
 * velocities and positions as updateMesh does. */
int mprgp_fem_begin_step(mprgp_fem *fem, mprgp_qp *qp, const mprgp_step_params *sp, int *rows);
int mprgp_fem_end_step(mprgp_fem *fem, mprgp_qp *qp, const mprgp_step_params *sp, double *delta);

/* ---- bandwidth probe ------------------------------------------------------
 * No counterpart in the reference: copies `bytes` between two device buffers
 * `iters` times with a streaming kernel and returns the HBM rate (read +
 * write bytes per second, GB/s) -- the achievable bound beside the published
 * peak, and the known-size kernel that calibrates profiler byte counters. */
int mprgp_copy_probe(mprgp_ctx *ctx, size_t bytes, int iters, double *gbps);

/* ---- serving engine (continuous batching) --------------------------------
 * No counterpart in the reference, which calls MPRGPDecoupledCon::solve once
 * per frame and per body (MprgpFemSolver.cpp:532-533); the engine is the
 * request loop around that call for many simulation sessions: a worker
 * thread keeps one batch of QPs advancing on the device and, between graph
 * replays, retires finished QPs and admits queued ones as slots and the HBM
 * budget allow.  Sessions keep their matrices and warm starts resident in
 * their QPs; a submitted QP is solved from its current x (the previous
 * solution unless set_x was called) exactly as mprgp_qp_solve would solve
 * it with the same tile shape: results are bit-identical alone or in any mix.
 */
typedef struct mprgp_engine mprgp_engine;
typedef struct mprgp_engine_config {
  int max_active;      /* QPs in the batch at once (0 = 256)                          */
  int tiles;           /* wavefronts per workgroup, 1/2/4/8 (0 = 4)                   */
  int graph_steps;     /* launches per captured graph replay (0 = 8)                  */
  int reserved_i[5];
  size_t hbm_budget;   /* bytes of matrices + QP vectors of admitted solves (0 = none) */
  double reserved_d[4];
} mprgp_engine_config;

typedef struct mprgp_engine_info {
  long long submitted, completed, rounds, launches;
  int active, queued, peak_active, reserved_i[5];
  size_t active_bytes, peak_bytes;
  double reserved_d[4];
} mprgp_engine_info;

void mprgp_engine_config_default(mprgp_engine_config *cfg);
int mprgp_engine_create(mprgp_ctx *ctx, const mprgp_engine_config *cfg, mprgp_engine **out);
/* Queues a solve of qp (rhs, rows and x set; any context on the engine's
 * device).  The QP is busy until its ticket is waited for.  MPRGP_EBUDGET if
 * this QP alone exceeds the budget. */
int mprgp_engine_submit(mprgp_engine *eng, mprgp_qp *qp, const mprgp_params *params, long long *ticket);
/* Waits up to timeout_ms (< 0: forever) for a ticket; returns its solve code
 * (0, -1, or an error) and fills stats (elapsed_ms = submit-to-completion
 * latency, queue_ms = queue wait, launches = launches while admitted), or MPRGP_EPENDING on time-out.  A ticket is waited for once. */
int mprgp_engine_wait(mprgp_engine *eng, long long ticket, int timeout_ms, mprgp_stats *stats);
/* As mprgp_engine_wait for whichever ticket completed first (in completion
 * order); *ticket receives it.  MPRGP_ESTATE if no ticket is outstanding. */
int mprgp_engine_wait_any(mprgp_engine *eng, int timeout_ms, long long *ticket, mprgp_stats *stats);
int mprgp_engine_get_info(mprgp_engine *eng, mprgp_engine_info *info);
/* Finishes every submitted solve, then stops the worker. */
void mprgp_engine_destroy(mprgp_engine *eng);

#ifdef __cplusplus
}
#endif
#endif /* MPRGP_AMD_H */
