// mprgp_dev.h -- device-side data layout shared by the HIP kernels and the
// host driver of the MI355X MPRGP contact-QP solver.
//
// Layout in HBM (per QP, all in the QP's internal node order, which sorts
// nodes by descending block-row length so that SELL slices carry no padding):
//   matrix  SELL-64 of 3x3 blocks: slice s covers internal nodes 64s..64s+63;
//           its blocks occupy [sl_off[s], sl_off[s+1]) (a multiple of 64);
//           block k of lane l sits at sl_off[s] + 64k + l; column index col[]
//           (internal node; padding slots name the row itself, or the last
//           node past the end, with zero values); values val[9*blk0 + 64*(9k+e) + l]
//           for entry e = 3*row + col of the block, so every load of one entry
//           across a wavefront is 64 consecutive values (512 B for FP64).
//           The mixed form keeps an FP32 copy of the values in the same layout.
//   vectors 3 values per node, node-major (x0 y0 z0 x1 ...).
//   rows    at most one half-space row per node: normal cn[3r..3r+2], offset
//           cc[r], |n|^2 in cj[r], tangent-Jacobi factor cs[r], flag cf[r]
//           (1 = node has a row), active set face[r].
#pragma once
#include <cstdint>

namespace mprgp {

constexpr int kSlice = 64;       // rows per SELL slice = lanes per wavefront
constexpr int kNumPart = 8;      // partial sums per tile (most any op produces: 7)
constexpr double kEps = 1e-9;    // ScalarUtil<double>::scalar_eps of the simulator
constexpr double kDblMax = 1.7976931348623157e308;  // ScalarUtil<double>::scalar_max

enum Op : int {
  OP_DONE = 0,
  OP_PROJX,      // x = P(x) before the solve (J/c facade overload)
  OP_POW_IT,     // power iteration k: v_k = w_{k-1}/nu_{k-1}, |v_k - v_{k-1}|^2,
                 // w_k = (A w_{k-1})/nu_{k-1} = A v_k, |w_k|^2
  OP_POW_SAVE,   // ev = v
  OP_INIT,       // g = Ax - b, face(x), phi, z, p = z, and the loop head (TOP)
  OP_TOP,        // loop head: phi, beta, |phi+beta|^2, beta.beta, phitphi
  OP_CG_AP,      // AP = A p, p.AP, z.g
  OP_CG_PY,      // y = x - alpha p, py = P(y), feasibility of y
  OP_CG_Y,       // A py (and A x at CG entry): objective values
  OP_CG_A,       // (x = y by buffer swap) g -= alpha AP, phi, z, z.AP, z.g,
                 // beta, |phi+beta|^2, beta.beta, phitphi
  OP_EXP1,       // y = P(x)
  OP_EXP2,       // g = A y - b, face(y), phi, y' = y - abar phi, x = P(y')
  OP_EXP3,       // g = A x - b, face(x), phi, z, p = z, y = y', and the loop head
  OP_PROP1,      // g = A x - b, AD = A beta, beta.AD, g.beta
  OP_PROP2,      // x = P(x - a beta), g -= a AD, face(x), phi, z, p = z, and
                 // the loop head
  OP_FINAL,      // x = P(x) (monotone only), active count
  OP_CL_AP,      // classic: AP = A p, p.AP, z.g, max feasible step
  OP_CL_CG,      // classic CG step: x -= a p, g -= a AP, phi, z, z.AP
  OP_CL_EXP,     // classic expansion (first half, elementwise)
  OP_CL_EXP2,    // classic expansion (second half, SpMV)
  OP_CL_PROP1,   // classic proportioning: AD = A beta
  OP_R_START,    // mixed: FP64 g = A x - b, FP64 KKT norm, correction QP set-up
  OP_R_UPD,      // mixed: x += d (FP64)
  OP_R_FINAL,    // mixed: x = P(x) in FP64, active count
  OP_CG_PYG,     // mixed inner CG step in one pass: y = x - alpha p,
                 // feasibility, p.g and f(P(y)) - f(y), and the gradient
                 // update of an accepted step applied in place: g -= alpha AP,
                 // phi, z, z.AP, z.g; then beta into the other beta buffer,
                 // |phi+beta|^2, beta.beta, phitphi
  OP_CG_BAP,     // p' = z - beta p into the other p buffer and AP = A p' with
                 // p' formed at the gather from z and p, p'.AP
  OP_CG_YA,      // FP64, after the first CG step: CG_Y, then CG_A applied
                 // speculatively (beta into the other beta buffer)
  OP_NUM
};

enum Variant : int { VAR_MONOTONIC = 0, VAR_CLASSIC = 1 };

// Per-QP control block.  Two copies per QP alternate with the launch parity:
// launch k reads copy (k+1)%2 (the state the previous launch executed) and
// the previous launch's per-tile partial sums, applies the control
// transition in every tile (identically), executes the resulting op, and its
// first tile stores the new state into copy k%2.  No device-scope fence is
// needed: launch boundaries order everything.
struct alignas(128) QpState {
  int op;
  int fresh;         // set by the host: no previous launch to reduce
  int n_tiles;
  int iteration, num_cg, num_exp, num_prop;
  int max_it;
  int prop;          // result_is_prop
  int cg_entry;      // next CG_Y also evaluates f(x)
  int y_valid, y_feas;
  int code;
  int active;
  int power_it;
  int ev_valid;      // ev holds a power vector on exit
  int has_rows;      // QP has at least one constraint row
  int precond;       // tangent-Jacobi preconditioning on
  int variant;
  int p_pending;     // classic: p = z - beta p still to apply
  int pow_next;      // op after the power method
  int steps;         // launches this QP took part in
  int mixed;         // inner ops run on the FP32 copies (mixed form)
  int refine, max_refine;   // mixed: outer sweeps done / allowed
  int iter_total;           // mixed: inner iterations over all sweeps
  int num_cg_t, num_exp_t, num_prop_t;
  int since_best;           // mixed inner: iterations without halving the residual
  int xb;                   // 1: x lives in the y buffer and y in the x buffer
  int pb;                   // 1: p lives in the p2 buffer (mixed inner)
  int bb;                   // 1: beta lives in the beta2 buffer (mixed inner)
  int prop_next;            // proportioning test at the CG trial point (mixed inner)
  int fp64_finish;          // mixed: the refinement ran out of sweeps; the FP64 form finishes
  double tol, alpha_bar, gamma;
  double alpha, beta_cg, pd, zg, residual, fx, fy;
  double alpha_f, nu, power_eps, lambda_max;
  double tol_outer, res64, inner_rel, best_res;
  double pg, fx_entry;      // p.g of the CG step; f(x) at CG entry (mixed inner)
  double zap;               // z.AP after the in-place gradient update (mixed inner)
  double res_next;          // KKT residual at the CG trial point (mixed inner)
  double fxu;               // f(x) of the iterate, unprojected, from the step sums (mixed inner)
  int opcnt[OP_NUM];  // launches this QP spent in each op
};
// The launch prologue moves a state copy as one 8-byte word per lane.
static_assert(sizeof(QpState) == 8 * kSlice, "QpState must stay 512 bytes");

// Storage of one precision form: FP64 for the plain solve and the outer
// refinement; for the inner solve of the mixed form, the matrix values and
// the gradient- and direction-side vectors (b, g, phi, beta, p, z, A p) are
// FP32.  The iterate and trial points (x, y, P(y)), the per-node rows and
// 1/diag stay FP64: the active set is decided on x against the 1e-9 plane
// tolerance, and the tangent-Jacobi factor s is computed from the FP64
// 1/diag, so the preconditioner must use that same 1/diag or it leaves a
// normal component eps32*|lambda| in z that pushes every CG step through the
// planes.  FP32 g itself carries eps32*|lambda| into the tangential part: the
// inner solve's KKT floor, far below the refinement's tolerance unless the
// contact forces lambda reach tol/eps32.
template <class T>
struct Vecs {
  const T *val;               // SELL values
  const double *dinv;         // 1/diag(A)
  double *x, *y, *py;         // iterate and trial points
  T *b, *g, *phi, *beta;      // gradient side
  T *p, *z, *ap;              // search direction side
  T *p2;                      // second p buffer (mixed inner: p' = z - beta p
                              // is written while neighbours still gather p)
  T *beta2;                   // second beta buffer (mixed inner: the trial
                              // point's beta, kept apart until it is accepted)
  const double *cn, *cj, *cs; // normal, |n|^2, tangent-Jacobi factor
  double *cc;                 // offset (mixed inner: shifted c - n.x)
};

struct QpDev {
  int n_nodes;
  int n_slices;
  const int *sl_off;
  const int *col;
  const unsigned short *col16;  // the same columns in 16 bits, or null
  Vecs<double> d;
  Vecs<float> f;
  double *ev;            // power vector kept for the next solve (FP64)
  double *pv, *pw;       // power iterates v_k by parity of k
  double *pu, *pz;       // products w_k = A v_k by parity of k
  const unsigned char *cf;
  unsigned char *face;
  int *done_flag;       // receives `tag` once the QP reaches DONE
  int tag;              // nonzero; distinguishes successive occupants of a batch slot
  const double *x0;     // non-null: restart_x copies it into d.x before the solve
};

// One application of the projector or preconditioner (projector_op).
// Variants of the step kernel; all give the same results, bit for bit.
//   kStepCached: SpMV unroll 3, matrices read through the caches;
//   kStepStream: the same with matrix values and column indices loaded
//                nontemporal (for launches whose matrices exceed the cache);
//   kStepTail:   SpMV unroll 6 and fewer waves (for launches with few tiles).
enum StepVariant { kStepCached = 0, kStepStream = 1, kStepTail = 2 };

enum ProjOp { PJ_PROJECT, PJ_FEASIBLE, PJ_FACE, PJ_PHI, PJ_BETA, PJ_PHITPHI, PJ_STEP_LIMIT, PJ_PRECOND };

struct ProjArgs {
  int op, n_nodes;
  const unsigned char *cf;
  const double *cn, *cc, *cj, *cs, *dinv;
  unsigned char *face;  // written by PJ_FACE, read by PJ_PHI, PJ_BETA, PJ_PRECOND
  const double *a, *b;  // x or g; phi or d
  double *out;          // per-node result
  double s0;            // alpha_bar (PJ_PHITPHI); rows present (PJ_PRECOND, 0 or 1)
  double *part;         // one partial per workgroup (ops that reduce), else null
};

// One work item of a launch: up to `slices` consecutive slices of one QP.
// The tile carries its QP's state and partial-sum pointers itself, so the
// launch prologue reaches them with one dependent load instead of going
// through the QP table.
struct Tile {
  QpState *st;   // the QP's st[0], st[1] (launch parity)
  double *part;  // the QP's [2][kNumPart][nt] partial sums (launch parity, slot-major)
  int qp;        // index into the batch's QpDev table
  int slice0;
  int slices;
  int local;     // tile index inside its QP (partials column)
  int nt;        // tiles of the QP
  int pad;
};

}  // namespace mprgp
