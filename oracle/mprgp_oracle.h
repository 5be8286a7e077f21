/*
 * mprgp_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU FP64 restatement of the reference's decoupled-constraint MPRGP QP solve
 * and of the FEM system assembly that produces its matrix.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or as the timed CPU baseline.  The product library never links
 * or calls anything here.
 *
 * Reference files followed (paths relative to the reference root):
 *   src/MPRGPSolver/MPRGPSolver.h       MPRGPBase / MPRGP / MPRGPMonotonic / FixedSparseMatrix / MPRGPDecoupledCon
 *   src/MPRGPSolver/MPRGPProjection.h   BaseGeneralConProjector / DecoupledConProjector
 *   src/MPRGPSolver/MPRGPPrecondition.h DiagonalInFacePreconSolver / DiagonalDecouplePreconSolver
 *   src/MPRGPSolver/MPRGPUtility.h      COUNT_CONSTRAINTS, getDiagonal
 *   src/ElasticSolver/FEMEnergy.cpp, FEMUtils.h, FEMMesh.cpp, FEMRotationUtil.h, FEMSystem.cpp
 *   src/CollisionHandle/MprgpFemSolver.cpp
 *
 * Pin status of every function is stated in the header comment of
 * mprgp_oracle.c / fem_oracle.c and checked by tests/test_oracle_*.py.
 */
#ifndef MPRGP_ORACLE_H
#define MPRGP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Compressed sparse rows.  For the system matrix A (square, symmetric) the
 * column indices of each row are sorted ascending, which makes the row-wise
 * dot product add terms in the same order as the reference's column-major
 * scatter SpMV over a symmetric matrix. */
typedef struct {
  int rows, cols;
  const int *ptr;
  const int *idx;
  const double *val;
} orc_csr;

typedef struct {
  int code;          /* 0 converged, -1 not converged (reference return code) */
  int iterations;    /* the reference's `iteration` counter ("total iter")      */
  int num_cg, num_exp, num_prop;
  int active;        /* COUNT_CONSTRAINTS(face) at exit                        */
  int power_iters;   /* power-method iterations (0 when alpha_bar was given)   */
  double residual;   /* residual_out = ||phi + beta||_2                        */
  double alpha_bar;
  double lambda_max; /* spectral-radius estimate used for alpha_bar            */
} orc_stats;

/* ---- matrix helpers -------------------------------------------------- */
void orc_spmv(const orc_csr *A, const double *x, double *y);
double orc_func_value(const orc_csr *A, const double *x, const double *b);
void orc_get_diagonal(const orc_csr *A, double *diag);

/* ---- projector (decoupled J: rows with disjoint column support) ------- */
typedef struct orc_projector orc_projector;
orc_projector *orc_projector_new(const orc_csr *J, const double *c);
void orc_projector_free(orc_projector *P);
int orc_projector_is_decoupled(const orc_projector *P);
void orc_project(const orc_projector *P, const double *x, double *y);
void orc_decide_face(orc_projector *P, const double *x);
int orc_is_feasible(const orc_projector *P, const double *x);
void orc_phi(const orc_projector *P, const double *g, double *phi);
void orc_beta(const orc_projector *P, const double *g, const double *phi, double *beta);
double orc_phitphi(const orc_projector *P, const double *x, double alpha_bar, const double *phi);
double orc_step_limit(const orc_projector *P, const double *x, const double *d, double alpha_cg);
const char *orc_face(const orc_projector *P);
int orc_num_active(const orc_projector *P);
void orc_set_face(orc_projector *P, const char *face);

/* ---- preconditioner ---------------------------------------------------- */
typedef struct orc_precond orc_precond;
orc_precond *orc_precond_new(const orc_csr *A, const orc_projector *P, int preconditioned);
void orc_precond_free(orc_precond *pc);
void orc_precond_solve(const orc_precond *pc, const orc_projector *P, const double *g,
                       double *z, const double *phi);

/* ---- spectral radius (power method) ----------------------------------- */
/* ev: in/out start vector of length n; *ev_valid says whether it holds one.
 * When no start vector is given, a deterministic pseudo-random one from
 * `seed` replaces the reference's setRandom(). */
double orc_spec_rad(const orc_csr *A, double *ev, int *ev_valid, double eps,
                    unsigned long long seed, int *iters_out);

/* ---- whole solves ------------------------------------------------------ */
enum { ORC_MONOTONIC = 0, ORC_CLASSIC = 1 };

/* MPRGPDecoupledCon<double>::solve<FixedSparseMatrix,preconditioned>(A,B,J,c,x,tol,max_it,name,ev)
 * project_first=1 reproduces the J/c overload (x is projected before the
 * solve); 0 reproduces the projector overload.  alpha_bar > 0 is used as is,
 * otherwise 2/specRad(A, ev, tol) as in the reference. */
int orc_decoupled_solve(const orc_csr *A, const double *B, const orc_csr *J, const double *c,
                        double *x, double tol, int max_it, int variant, int preconditioned,
                        double alpha_bar, double *ev, int *ev_valid, int project_first,
                        orc_stats *st);

/* ---- traced solves ------------------------------------------------------ */
/* One decision of the monotone solve, in the order the control flow takes
 * them (iteration = the reference's `iteration` counter at that point):
 *   ORC_EV_RESIDUAL  residual_out <= tol at the loop head or after a CG
 *                    step (lhs = residual, rhs = tol, outcome = converged)
 *   ORC_EV_PROP      proportional(x) (lhs = beta.beta, rhs = gamma^2 phitphi)
 *   ORC_EV_CG_TEST   fy <= fx of the monotone CG loop (lhs = fy, rhs = fx,
 *                    outcome = step accepted); when y is infeasible the next
 *                    event is ORC_EV_Y_INFEASIBLE (lhs = min over rows of
 *                    n.y - c)
 *   ORC_EV_FEASIBLE  isFeasible(y) for the choice expansion/proportioning
 *                    (lhs = min over rows of n.y - c, rhs = -1e-9)
 *   ORC_EV_STEP      the step taken (outcome ORC_STEP_*)
 *   ORC_EV_FY_CLOSED follows each ORC_EV_CG_TEST: f(P(y)) of the same trial
 *                    point by the closed forms of ORC_F_CLOSED (lhs) and
 *                    ORC_F_CLOSED_FROM_FPY (rhs), each tracked along this
 *                    solve's own trajectory */
enum { ORC_EV_RESIDUAL = 0, ORC_EV_PROP = 1, ORC_EV_CG_TEST = 2, ORC_EV_FEASIBLE = 3, ORC_EV_STEP = 4,
       ORC_EV_Y_INFEASIBLE = 5, ORC_EV_FY_CLOSED = 6 };
enum { ORC_STEP_CG = 0, ORC_STEP_EXP = 1, ORC_STEP_PROP = 2 };
typedef struct {
  int iteration, kind, outcome, pad;
  double lhs, rhs;
} orc_event;
/* How the monotone CG loop obtains f(P(y)) (MPRGPSolver.h:283-308):
 *   ORC_F_DIRECT  as the reference: project y, then 1/2 y'Ay - y'b by a SpMV;
 *   ORC_F_CLOSED  from the step's sums, as the device's mixed form does:
 *                 f(y) = f(x) - alpha p'g + alpha^2 p'Ap/2, with f(x) the
 *                 objective at the (unprojected) iterate, plus
 *                 d'(A y - b) + d'A d/2 for d = P(y) - y taken on the rows
 *                 violated by more than the feasibility tolerance 1e-9;
 *   ORC_F_CLOSED_FROM_FPY  the same, but continuing from f(P(y)) of the
 *                 accepted step instead of f(y): a deliberately wrong form
 *                 that tests use to show the checks catch it. */
enum { ORC_F_DIRECT = 0, ORC_F_CLOSED = 1, ORC_F_CLOSED_FROM_FPY = 2 };
typedef struct {
  orc_event *events; /* may be NULL */
  int cap;           /* capacity of events                                  */
  int n;             /* out: events recorded (may exceed cap: the rest lost) */
  int fmode;         /* ORC_F_*                                              */
} orc_options;
/* orc_decoupled_solve with options (the monotone variant records events). */
int orc_decoupled_solve_ex(const orc_csr *A, const double *B, const orc_csr *J, const double *c,
                           double *x, double tol, int max_it, int variant, int preconditioned,
                           double alpha_bar, double *ev, int *ev_valid, int project_first,
                           orc_options *opt, orc_stats *st);

/* ---- coupled rows (general_oracle.c) ----------------------------------- */
/* GeneralConProjector (MPRGPProjection.h:400-520) for rows spanning several
 * nodes, and MPRGPGeneralCon<double>::solve (MPRGPSolver.h:594-626). */
typedef struct orc_gprojector orc_gprojector;
/* exact argmin 1/2 l'Gl - h'l s.t. l >= 0 (G m x m row-major SPD); 0 or -1 */
int orc_qp_nonneg(int m, const double *G, const double *h, double *lam);
orc_gprojector *orc_gprojector_new(const orc_csr *J, const double *c);
void orc_gprojector_free(orc_gprojector *P);
int orc_gproject(const orc_gprojector *P, const double *x, double *y);
int orc_gdecide_face(orc_gprojector *P, const double *x);
int orc_gis_feasible(const orc_gprojector *P, const double *x);
void orc_gphi(const orc_gprojector *P, const double *g, double *phi);
int orc_gbeta(const orc_gprojector *P, const double *g, const double *phi, double *beta);
double orc_gphitphi(const orc_gprojector *P, const double *x, double alpha_bar, const double *phi);
const char *orc_gface(const orc_gprojector *P);
int orc_gnum_active(const orc_gprojector *P);
/* J/c overload: x projected first when project_first; the monotone variant;
 * preconditioned = DiagonalInFacePreconSolver (z = phi D^-1), else z = phi.
 * Returns 0, -1 (not converged) or -2 (a projection QP failed). */
int orc_general_solve(const orc_csr *A, const double *B, const orc_csr *J, const double *c, double *x, double tol,
                      int max_it, int preconditioned, double alpha_bar, double *ev, int *ev_valid,
                      int project_first, orc_stats *st);

/* ---- ICA comparator (ica_oracle.c) ------------------------------------ */
/* ICASolver (ICASolver.h:287-463) with its ProjectedGaussSeidel (:20-150). */
typedef struct {
  int code;           /* 0: |A x - R - J^T lambda| <= tol reached, -1 not        */
  int iterations;     /* the reference's `iterations` (index at the stop, or max) */
  int gs_iterations;  /* projected Gauss-Seidel sweeps run, summed over calls    */
  double residual;    /* |A x - R - J^T lambda| at exit                          */
  double gs_residual; /* the last PGS call's |lambda - lambda_prev|              */
} orc_ica_stats;
void orc_inv3(const double *m, double *inv);
int orc_pgs(const orc_csr *B, const double *c, double *lam, int max_it, double tol, int *iters, double *res);
/* B = J Da^-1 J^T as CSR (arrays malloc'd; free with orc_ica_free) */
int orc_ica_b(const orc_csr *J, int n_nodes, const double *dinv, int *nnz_out, int **ptr, int **idx, double **val);
void orc_ica_free(void *p);
int orc_ica_solve(const orc_csr *A, const double *R, const orc_csr *J, const double *p, double *x, double tol,
                  int max_it, int gs_max_it, double gs_tol, double *lambda, orc_ica_stats *st);

/* ---- FEM assembly (fem_oracle.c) -------------------------------------- */
enum { ORC_LINEAR = 1, ORC_COROTATIONAL = 2 };

/* FEMBody::assemble + makePositive + buildD: orients tets positively
 * (tets is modified in place), fills d (9 per tet, row-major Dm^-1),
 * vol (per tet) and node_mass (per node, sum V/4, not yet times density). */
void orc_fem_setup(int n_nodes, int n_tets, int *tets, const double *X0,
                   double *d, double *vol, double *node_mass);

/* Polar rotation R = U V^T following adjustF3D (eigen-decomposition of F^T F). */
void orc_polar_rotation(const double *F, double *R);

/* MaterialEnergy::eval: element force fB (12) and stiffness HB (12x12,
 * row-major) of one tet at current node positions xe (4x3, node-major). */
double orc_fem_element(int type, const double *xe, const double *d, double vol,
                       double lambda, double mu, double *fB, double *HB);

/* MprgpFemSolver::buildLinearSystem for one body, first Newton iteration
 * (x1 = x_n): A = M + h^2 K, b = h^2 f(x1) + h^2 M g + M (x_n + h v_n - x1) + A x1.
 * young/poisson are per tet.  Returns the CSR in caller-freed arrays
 * (orc_free).  Returns 0 on success. */
int orc_fem_build_system(int n_nodes, int n_tets, const int *tets_oriented,
                         const double *d, const double *vol, const double *node_mass,
                         const double *young, const double *poisson, double density,
                         const double *x_n, const double *v_n, const double *gravity,
                         double h, int type,
                         int **ptr_out, int **idx_out, double **val_out, double *b_out);
/* The same at a Newton iterate x1 (buildLinearSystem after updatePos,
 * MprgpFemSolver.cpp:315-356): elements evaluated at x1,
 * b = h^2 f(x1) + h^2 M g + M (x_n + h v_n - x1) + A x1. */
int orc_fem_build_system_at(int n_nodes, int n_tets, const int *tets_oriented,
                            const double *d, const double *vol, const double *node_mass,
                            const double *young, const double *poisson, double density,
                            const double *x_n, const double *v_n, const double *x1,
                            const double *gravity, double h, int type,
                            int **ptr_out, int **idx_out, double **val_out, double *b_out);
/* The same with Rayleigh damping alpha_k (mass) and alpha_m (stiffness)
 * (FEMSystem::setDamping as Simulator.cpp:159-167 sets it): A = M (1 + h
 * alpha_k) + K (h^2 + h alpha_m), b += alpha_m K c + M c alpha_k with
 * c = (x_n - x1) h. */
int orc_fem_build_system_damped(int n_nodes, int n_tets, const int *tets_oriented,
                                const double *d, const double *vol, const double *node_mass,
                                const double *young, const double *poisson, double density,
                                const double *x_n, const double *v_n, const double *x1,
                                const double *gravity, double h, int type, double alpha_k,
                                double alpha_m, int **ptr_out, int **idx_out, double **val_out,
                                double *b_out);
/* MprgpFemSolver::updateMesh (MprgpFemSolver.cpp:377-390) with the implicit
 * Euler parameters of FEMSolver::resetImplicitEuler (FEMSystem.h:140-142). */
void orc_implicit_euler_velocity(int n, const double *x_n, const double *v_n, const double *x,
                                 double h, double *v);
void orc_free(void *p);

/* ---- plane collider (collide_oracle.c) --------------------------------- */
/* Decoupled discrete collision of one body's n_nodes positions x (3 per node)
 * with n_planes planes {nx, ny, nz, p, delta} in order: the rows of
 * LinearConCollider::getConstraints after handle(body, plane, delta) for each
 * plane.  node/normal/c receive at most n_nodes rows (normal 3 per row).
 * Returns the row count, or -1 when a normal's length is outside
 * [1 - 1e-8, 1 + 1e-8]. */
int orc_plane_collide(int n_nodes, const double *x, int n_planes, const double *planes,
                      int *node, double *normal, double *c);

#ifdef __cplusplus
}
#endif
#endif
