// mprgp_kernels.hip -- CDNA4 (gfx950) kernels of the MPRGP contact-QP solve.
//
// One kernel, `mprgp_step`, advances every QP of a batch by one operation of
// the reference's MPRGP control flow (MPRGPSolver.h:148-456).  Each QP keeps
// its own state machine on the device (QpState::op).  A launch first applies,
// in every tile, the control transition of the op the QP executed in the
// previous launch (from that launch's per-tile partial sums, reduced in a
// fixed order), then executes the resulting op on the tile's rows.  The host
// never reads a scalar of the iteration: it replays launches (captured in a
// hipGraph) and polls a done-counter.
//
// Work mapping: one wavefront per SELL slice, one lane per block row
// (= node).  SpMV ops stream the slice's 3x3 blocks with fully coalesced
// loads (one block entry of 64 consecutive rows is contiguous) and gather
// the input vector per neighbour node.  Elementwise ops (projection, gradient
// split, preconditioner, vector updates) run in the same lane-per-node
// mapping and are fused into the SpMV ops wherever the data flow allows.
//
// Precision: arithmetic and every reduction run in FP64.  The plain form
// stores matrix and vectors in FP64; the mixed form runs the inner solve on
// an FP32 copy of the matrix (and FP32 gradient- and direction-side vectors;
// iterates, rows and 1/diag stay FP64) inside an FP64 iterative refinement:
// the residual b - A x, the KKT norm and the update x += d use the FP64 data.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mprgp_host.h"

namespace mprgp {


struct V3 {
  double x, y, z;
};

// The data pointers of the QP table are generic pointers; every access
// through them goes via gmem() to the global address space, so it compiles to
// global memory instructions rather than flat ones (which would also count
// against the LDS/scalar wait counter and take 64-bit VGPR addresses).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

template <class T>
__device__ __forceinline__ V3 ld3(const T *v, int r) {
  const auto *p = gmem(v) + 3 * (size_t)r;
  return V3{(double)p[0], (double)p[1], (double)p[2]};
}
// One element; NT: nontemporal (streamed past the caches).
template <bool NT, class P>
__device__ __forceinline__ auto ldm(const P *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <class T>
__device__ __forceinline__ void st3(T *v, int r, V3 a) {
  auto *p = gmem(v) + 3 * (size_t)r;
  p[0] = (T)a.x;
  p[1] = (T)a.y;
  p[2] = (T)a.z;
}
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// a - s*b
__device__ __forceinline__ V3 axmy(V3 a, double s, V3 b) {
  return V3{a.x - s * b.x, a.y - s * b.y, a.z - s * b.z};
}
__device__ __forceinline__ double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The half-space row of one node (at most one per node).
struct Row {
  V3 n;
  double c, jj, s;
  bool has;
};

template <class T>
__device__ __forceinline__ Row ldrow(const QpDev &q, const Vecs<T> &v, int r) {
  Row R;
  R.has = gmem(q.cf)[r] != 0;
  if (R.has) {
    R.n = ld3(v.cn, r);
    R.c = gmem(v.cc)[r];
    R.jj = gmem(v.cj)[r];
    R.s = gmem(v.cs)[r];
  } else {
    R.n = V3{0.0, 0.0, 0.0};
    R.c = R.jj = R.s = 0.0;
  }
  return R;
}

// DecoupledConProjector::project (MPRGPProjection.h:333-348), one node.
__device__ __forceinline__ V3 proj(const Row &R, V3 x) {
  if (!R.has) return x;
  double v = (R.c - dot3(R.n, x)) / R.jj;
  double lam = (v < 0.0) ? 0.0 : v;
  return V3{x.x + R.n.x * lam, x.y + R.n.y * lam, x.z + R.n.z * lam};
}

// DECIDE_FACE (MPRGPProjection.h:283-297): |n.x - c| < eps.
__device__ __forceinline__ bool face_of(const Row &R, V3 x) {
  return R.has && fabs(dot3(R.n, x) - R.c) < kEps;
}

// isFeasible (MPRGPProjection.h:298-310): violated when n.x < c - eps.
__device__ __forceinline__ double infeasible(const Row &R, V3 x) {
  return (R.has && dot3(R.n, x) < R.c - kEps) ? 1.0 : 0.0;
}

// PHI (MPRGPProjection.h:349-367): free gradient.
__device__ __forceinline__ V3 phi_of(const Row &R, bool face, V3 g) {
  if (!face) return g;
  double lam = -dot3(R.n, g) / R.jj;
  return V3{g.x + R.n.x * lam, g.y + R.n.y * lam, g.z + R.n.z * lam};
}

// BETA (MPRGPProjection.h:368-383): chopped gradient.  Rows outside the
// face subtract lambda = n.beta unscaled, exactly as the reference does.
__device__ __forceinline__ V3 beta_of(const Row &R, bool face, V3 g, V3 phi) {
  V3 b = sub(g, phi);
  if (!R.has) return b;
  double lam = dot3(R.n, b);
  if (face) {
    double v = lam / R.jj;
    lam = (0.0 < v) ? v : 0.0;
  }
  return V3{b.x - R.n.x * lam, b.y - R.n.y * lam, b.z - R.n.z * lam};
}

// DiagonalDecouplePreconSolver::solve (MPRGPPrecondition.h:91-105):
// z = D^-1 g - (n o D^-1) lambda with lambda = s (n o D^-1).g on active rows;
// with preconditioning off or no rows at all, z = phi.
template <class T>
__device__ __forceinline__ V3 precond_of(const Vecs<T> &v, bool on, const Row &R, bool face, int r, V3 g,
                                         V3 phi) {
  if (!on) return phi;
  V3 d = ld3(v.dinv, r);
  V3 z{d.x * g.x, d.y * g.y, d.z * g.z};
  if (face) {
    V3 jm{R.n.x * d.x, R.n.y * d.y, R.n.z * d.z};
    double lam = R.s * dot3(jm, g);
    z = V3{z.x - jm.x * lam, z.y - jm.y * lam, z.z - jm.z * lam};
  }
  return z;
}

// ------------------------------------------------------------------ SpMV

// Accumulates A v for the lane's block row over the slice's blocks; `gather`
// returns NV input vectors at a neighbour node.  Padding slots hold a valid
// column and zero values, so the loop has no branch and the column indices,
// matrix values and gathers of U blocks are all in flight together.  U is
// kept small enough (3 blocks for one vector, 1 for two) that the step kernel
// needs no scratch: spilled registers would go to HBM on every launch.
// NT: matrix values and column indices are loaded nontemporal.  Each is read
// once per launch and by one lane only; when the launch's matrices are larger
// than the last-level cache, streaming them past it keeps the gathered
// vectors cached (measured on 1.8 GB of FP64 matrices: batched SpMV 4.70 ->
// 4.97 TB/s).  Matrices that fit stay cached from launch to launch instead.
// C: the column index type, 32-bit or 16-bit (2 of the 40 bytes an FP32
// block streams: on the 256-QP batch, 3.1% shorter launches).
template <int NV, int U = (NV == 1 ? 3 : 1), bool NT = false, class C, class T, class G>
__device__ __forceinline__ void spmv_row(const int *sl_off, const C *col, const T *val, int s, int lane,
                                         G gather, V3 (&acc)[NV]) {
#pragma unroll
  for (int t = 0; t < NV; ++t) acc[t] = V3{0.0, 0.0, 0.0};
  const int b0 = gmem(sl_off)[s];
  const int W = (gmem(sl_off)[s + 1] - b0) / kSlice;
  const auto *cp = gmem(col) + b0 + lane;
  const auto *vp = gmem(val) + (size_t)b0 * 9 + lane;
  int k = 0;
  for (; k + U <= W; k += U) {
    int j[U];
    T a[U][9];  // kept in the storage precision until its multiply
    V3 xs[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = (int)ldm<NT>(cp + (k + u) * kSlice);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < 9; ++e) a[u][e] = ldm<NT>(vp + (size_t)(k + u) * (9 * kSlice) + e * kSlice);
#pragma unroll
    for (int u = 0; u < U; ++u) gather(j[u], xs[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int t = 0; t < NV; ++t) {
        acc[t].x += (double)a[u][0] * xs[u][t].x;
        acc[t].x += (double)a[u][1] * xs[u][t].y;
        acc[t].x += (double)a[u][2] * xs[u][t].z;
        acc[t].y += (double)a[u][3] * xs[u][t].x;
        acc[t].y += (double)a[u][4] * xs[u][t].y;
        acc[t].y += (double)a[u][5] * xs[u][t].z;
        acc[t].z += (double)a[u][6] * xs[u][t].x;
        acc[t].z += (double)a[u][7] * xs[u][t].y;
        acc[t].z += (double)a[u][8] * xs[u][t].z;
      }
  }
  for (; k < W; ++k) {
    const int j = (int)ldm<NT>(cp + k * kSlice);
    T a[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] = ldm<NT>(vp + (size_t)k * (9 * kSlice) + e * kSlice);
    V3 xs[NV];
    gather(j, xs);
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      acc[t].x += (double)a[0] * xs[t].x;
      acc[t].x += (double)a[1] * xs[t].y;
      acc[t].x += (double)a[2] * xs[t].z;
      acc[t].y += (double)a[3] * xs[t].x;
      acc[t].y += (double)a[4] * xs[t].y;
      acc[t].y += (double)a[5] * xs[t].z;
      acc[t].z += (double)a[6] * xs[t].x;
      acc[t].z += (double)a[7] * xs[t].y;
      acc[t].z += (double)a[8] * xs[t].z;
    }
  }
}

// ----------------------------------------------------------- reductions

// Cross-lane moves without the LDS crossbar: DPP within rows of 16 lanes,
// then readlane of the four row totals.
constexpr int kDppSwap1 = 0xb1;        // quad_perm [1,0,3,2]: lane ^ 1
constexpr int kDppSwap2 = 0x4e;        // quad_perm [2,3,0,1]: lane ^ 2
constexpr int kDppHalfMirror = 0x141;  // lane i of each 8 reads lane 7 - i
constexpr int kDppMirror = 0x140;      // lane i of each 16 reads lane 15 - i

template <int kCtrl>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the wavefront in a fixed order (pairs, quads, eighths, rows of
// 16, then the four rows), the same value in every lane: each combining
// step adds the same two operands in both lanes of a partner pair.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<kDppSwap1>(v);
  v += dpp<kDppSwap2>(v);
  v += dpp<kDppHalfMirror>(v);
  v += dpp<kDppMirror>(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ __forceinline__ double wave_min(double v) {
  v = fmin(v, dpp<kDppSwap1>(v));
  v = fmin(v, dpp<kDppSwap2>(v));
  v = fmin(v, dpp<kDppHalfMirror>(v));
  v = fmin(v, dpp<kDppMirror>(v));
  return fmin(fmin(lane_value(v, 0), lane_value(v, 16)), fmin(lane_value(v, 32), lane_value(v, 48)));
}

// Partial sums an op produces: slots 0..n-1 reduce with sum.  The classic
// step-length search (CL_AP) also reduces slot 3 with min.
constexpr int kMinSlot = 3;
__device__ __forceinline__ int sum_slots(int op) {
  switch (op) {
    case OP_CG_PYG:
      return 8;
    case OP_CG_YA:
      return 7;
    case OP_CG_A:
      return 5;
    case OP_TOP: case OP_CG_AP: case OP_INIT: case OP_EXP3: case OP_PROP2:
      return 3;
    case OP_CG_Y:
      return 4;
    case OP_CL_AP: case OP_PROP1: case OP_CL_PROP1: case OP_POW_IT:
      return 2;
    case OP_CG_PY: case OP_EXP2: case OP_FINAL: case OP_CL_CG: case OP_CG_BAP:
    case OP_R_START: case OP_R_FINAL:
      return 1;
    default:
      return 0;
  }
}

// ------------------------------------------------------------- control

struct Sums {
  double v[kNumPart];
};

// Op chosen after the outer loop's `iteration < max_iterations` test.
__device__ __forceinline__ void next_outer(QpState &st) {
  st.op = (st.iteration < st.max_it) ? OP_TOP : OP_FINAL;
}

// proportional (MPRGPSolver.h:174-181) with PHITPHI's scaling
// (MPRGPProjection.h:384-393).
__device__ __forceinline__ int proportional(const QpState &st, double bb, double xpx_phi) {
  double r = xpx_phi * (1.0 / st.alpha_bar);
  double phitphi = r > 0.0 ? r : 0.0;
  return bb <= st.gamma * st.gamma * phitphi;
}

// Mixed form only: the inner solve on FP32 data has a precision floor; when
// 128 iterations pass without halving the best residual the inner solve
// ends and the FP64 refinement takes over.  The FP64 form never stops early.
__device__ __forceinline__ bool inner_stalled(QpState &st) {
  if (!st.mixed) return false;
  if (st.residual < 0.5 * st.best_res) {
    st.best_res = st.residual;
    st.since_best = st.iteration;
    return false;
  }
  return st.iteration - st.since_best > 128;
}

// CGMonotonicStep's acceptance test (MPRGPSolver.h:284-292) once f(P(y)) is
// known: fnew = f(P(y)), fx_entry = f(x) when the CG loop is entered.
// applied: the trial launch already applied the accepted step's update
// (g -= alpha AP, phi, z in place, beta into the other buffer) and reduced
// its sums (CG_PYG, CG_YA); otherwise CG_A does that after acceptance.
__device__ __forceinline__ void cg_decide(QpState &st, double fnew, double fx_entry, bool applied) {
  if (st.cg_entry) {
    st.fx = fx_entry;
    st.fy = fnew;
    st.prop = 1;
    st.cg_entry = 0;
  } else {
    st.fx = st.fy;
    st.fy = fnew;
    st.num_cg++;
    st.iteration++;
  }
  if (st.prop && st.residual > st.tol && st.fy <= st.fx && st.iteration < st.max_it) {
    st.xb ^= 1;  // x = y: the trial point's buffer becomes the iterate's
    if (applied) {
      // g -= alpha AP, phi, z, beta and the KKT and proportioning sums at y
      // are done (MPRGPSolver.h:293-302)
      st.bb ^= 1;
      st.beta_cg = st.zap / st.pd;
      st.residual = st.res_next;
      st.prop = st.prop_next;
      // a stalled inner solve ends through the loop's exit path
      st.op = inner_stalled(st) ? OP_FINAL : OP_CG_BAP;
    } else {
      st.op = OP_CG_A;
    }
  } else if (st.residual <= st.tol) {
    st.op = OP_FINAL;
  } else {
    st.op = (st.prop || !st.y_valid || !st.y_feas) ? OP_EXP1 : OP_PROP1;
  }
}

__device__ __forceinline__ void finish_power(QpState &st) {
  st.lambda_max = st.nu;
  st.alpha_bar = 2.0 / st.nu;
  st.op = st.mixed ? OP_R_START : OP_INIT;
}

// The outer loop head (MPRGPSolver.h:417-437): KKT residual and
// proportioning from the sums of computeGradients + proportional (slots 0-2).
// TOP reduces them; INIT, EXP3 and PROP2 reduce them too, at the gradient
// they produce, so that no separate TOP launch follows them.
__device__ __forceinline__ void loop_head(QpState &st, const Sums &S) {
  st.p_pending = 0;
  st.residual = sqrt(S.v[0]);
  if (st.residual <= st.tol || inner_stalled(st)) {
    st.op = OP_FINAL;
    return;
  }
  st.prop = proportional(st, S.v[1], S.v[2]);
  if (st.variant == VAR_CLASSIC) {
    st.op = st.prop ? OP_CL_AP : OP_CL_PROP1;
  } else if (st.prop) {
    st.cg_entry = 1;
    st.op = OP_CG_AP;
  } else {
    st.op = (!st.y_valid || !st.y_feas) ? OP_EXP1 : OP_PROP1;
  }
}

// next_outer for an op that carried the loop head's sums.
__device__ __forceinline__ void next_outer_head(QpState &st, const Sums &S) {
  if (st.iteration < st.max_it) loop_head(st, S);
  else st.op = OP_FINAL;
}

// Consumes the reduced partial sums of the op a QP executed in the previous
// launch and moves its state machine to the next op (MPRGPSolver.h:42-79,
// 148-456).  Every tile of the QP runs it on identical inputs.
__device__ void control(QpState &st, const Sums &S) {
  switch (st.op) {
    case OP_PROJX:
      st.op = st.pow_next;
      break;
    case OP_POW_IT: {
      // specRad (MPRGPSolver.h:57-76) one launch per iteration: launch k
      // closes iteration k-1 (its delta test) and runs iteration k's product
      const int k = st.power_it;
      if (k >= 1) {
        if (sqrt(S.v[1]) <= st.power_eps) {  // converged at k-1: ev = v_{k-1}, nu_{k-1}
          st.power_it = k - 1;
          st.op = OP_POW_SAVE;
          break;
        }
        if (k > 1000) {  // iterations 0..1000 done: ev untouched
          finish_power(st);
          break;
        }
      }
      const double nu = sqrt(S.v[0]);
      st.nu = nu;
      if (nu < kEps) {  // tiny norm ends the iteration, ev = v_k
        st.nu = kEps;
        st.op = OP_POW_SAVE;
        break;
      }
      st.power_it = k + 1;
      st.op = OP_POW_IT;
      break;
    }
    case OP_POW_SAVE:
      st.ev_valid = 1;
      finish_power(st);
      break;
    case OP_INIT:
      st.iteration = st.num_cg = st.num_exp = st.num_prop = 0;
      st.residual = 0.0;
      st.y_valid = 0;
      st.best_res = kDblMax;
      st.since_best = 0;
      if (st.max_it > 0) loop_head(st, S);  // INIT carried the loop head's sums
      else st.op = OP_FINAL;
      break;
    case OP_TOP:
      loop_head(st, S);
      break;
    case OP_CG_AP:
      st.pd = S.v[0];
      st.zg = S.v[1];
      st.alpha = st.zg / st.pd;
      if (st.mixed && st.cg_entry) st.fx_entry = 0.5 * S.v[2];  // f(x) = x'(g - b)/2, g = A x - b
      st.op = st.mixed ? OP_CG_PYG : OP_CG_PY;
      break;
    case OP_CG_PY:
      st.y_valid = 1;
      st.y_feas = S.v[0] == 0.0;
      // the CG entry needs f(x) too (four sums): its step takes CG_Y, CG_A
      st.op = st.cg_entry ? OP_CG_Y : OP_CG_YA;
      break;
    case OP_CG_PYG: {
      // mixed inner solve: f(y) = f(x) - alpha p'g + alpha^2 p'Ap/2 and
      // f(P(y)) = f(y) + the projected rows' share, all from the step's own
      // sums.  f(x) is the objective at the iterate itself: at the CG entry
      // f(x) = x'(g - b)/2; after an accepted step x = y, unprojected
      // (MPRGPSolver.h:293), so f(x) is that step's f(y), not the f(P(y))
      // the acceptance test compares (fx = fy, :307).
      st.y_valid = 1;
      st.y_feas = S.v[0] == 0.0;
      st.pg = S.v[1];
      st.zg = S.v[2];  // z.g after the update: the next step's alpha numerator
      st.zap = S.v[3];
      st.res_next = sqrt(S.v[4]);
      st.prop_next = proportional(st, S.v[5], S.v[6]);
      const double f0 = st.cg_entry ? st.fx_entry : st.fxu;
      st.fxu = f0 - st.alpha * st.pg + 0.5 * st.alpha * st.alpha * st.pd;  // f(y): f(x) once accepted
      cg_decide(st, st.fxu + S.v[7], st.fx_entry, true);
      break;
    }
    case OP_CG_Y:
      cg_decide(st, 0.5 * S.v[0] - S.v[1], 0.5 * S.v[2] - S.v[3], false);
      break;
    case OP_CG_YA:
      st.zap = S.v[2];
      st.zg = S.v[3];  // z.g after the update: the next step's alpha numerator
      st.res_next = sqrt(S.v[4]);
      st.prop_next = proportional(st, S.v[5], S.v[6]);
      cg_decide(st, 0.5 * S.v[0] - S.v[1], 0.0, true);  // fx_entry: not the CG entry
      break;
    case OP_CG_A:
      st.beta_cg = S.v[0] / st.pd;
      st.zg = S.v[1];
      st.residual = sqrt(S.v[2]);
      st.prop = proportional(st, S.v[3], S.v[4]);
      st.op = OP_CG_BAP;
      break;
    case OP_CG_BAP:
      st.pb ^= 1;  // p' went to the other buffer
      st.pd = S.v[0];
      st.alpha = st.zg / st.pd;
      st.op = st.mixed ? OP_CG_PYG : OP_CG_PY;
      break;
    case OP_EXP1:
      st.op = OP_EXP2;
      break;
    case OP_EXP2:
      st.y_valid = 1;
      st.y_feas = S.v[0] == 0.0;
      st.op = OP_EXP3;
      break;
    case OP_EXP3:
      st.num_exp++;
      st.iteration++;
      next_outer_head(st, S);
      break;
    case OP_PROP1:
      st.alpha = S.v[1] / S.v[0];
      st.op = OP_PROP2;
      break;
    case OP_PROP2:
      st.num_prop++;
      st.iteration++;
      next_outer_head(st, S);
      break;
    case OP_CL_AP: {
      st.pd = S.v[0];
      st.zg = S.v[1];
      double acg = st.zg / st.pd;
      double af = (acg == kDblMax) ? acg : (acg + kEps);
      double lim = S.v[kMinSlot];
      if (lim < af) af = lim;
      st.alpha = acg;
      st.alpha_f = af;
      st.op = (acg <= af) ? OP_CL_CG : OP_CL_EXP;
      break;
    }
    case OP_CL_CG:
      st.beta_cg = S.v[0] / st.pd;
      st.p_pending = 1;
      st.num_cg++;
      st.iteration++;
      next_outer(st);
      break;
    case OP_CL_EXP:
      st.op = OP_CL_EXP2;
      break;
    case OP_CL_EXP2:
      st.num_exp++;
      st.iteration++;
      next_outer(st);
      break;
    case OP_CL_PROP1:
      st.alpha = S.v[1] / S.v[0];
      st.op = OP_PROP2;
      break;
    case OP_FINAL:
      if (st.mixed) {  // inner solve done: fold its counts, apply x += d
        st.iter_total += st.iteration;
        st.num_cg_t += st.num_cg;
        st.num_exp_t += st.num_exp;
        st.num_prop_t += st.num_prop;
        st.op = OP_R_UPD;
        break;
      }
      st.active = (int)S.v[0];
      st.code = st.residual <= st.tol ? 0 : -1;
      if (st.fp64_finish) {  // the counts of the mixed sweeps before it
        st.iteration += st.iter_total;
        st.num_cg += st.num_cg_t;
        st.num_exp += st.num_exp_t;
        st.num_prop += st.num_prop_t;
      }
      st.op = OP_DONE;
      break;
    case OP_R_START: {
      st.res64 = sqrt(S.v[0]);
      if (!(st.res64 == st.res64)) {  // NaN: stop
        st.op = OP_R_FINAL;
      } else if (st.res64 <= st.tol_outer) {
        st.op = OP_R_FINAL;
      } else if (st.refine >= st.max_refine) {
        // every sweep used and still above tol: the inner solve's FP32 floor
        // (about eps32 |lambda| per active row, the contact force leaking
        // into the tangential gradient) is above what the sweeps need.  The
        // FP64 form solves on from the current x (its own vectors, the rows'
        // own offsets, x in the x buffer).
        st.mixed = 0;
        st.fp64_finish = 1;
        st.tol = st.tol_outer;
        st.xb = st.pb = st.bb = 0;
        st.op = OP_INIT;
      } else {
        double t = st.inner_rel * st.res64;
        st.tol = t > 0.5 * st.tol_outer ? t : 0.5 * st.tol_outer;
        st.xb = 0;  // the correction d = 0 was written to the first buffer
        st.op = OP_INIT;
      }
      break;
    }
    case OP_R_UPD:
      st.refine++;
      st.op = OP_R_START;
      break;
    case OP_R_FINAL:
      st.active = (int)S.v[0];
      st.residual = st.res64;
      st.code = st.res64 <= st.tol_outer ? 0 : -1;
      st.iteration = st.iter_total;
      st.num_cg = st.num_cg_t;
      st.num_exp = st.num_exp_t;
      st.num_prop = st.num_prop_t;
      st.op = OP_DONE;
      break;
    default:
      break;
  }
}

// --------------------------------------------------------------- kernel

// One row (= node r of slice s) of a solve op on storage T; partial sums
// accumulate into P (FP64).
// MX: the kernel instance may hold mixed-form QPs (otherwise their branches
// are compiled out).
template <class T>
__device__ __forceinline__ V3 round_to(V3 a) {  // the value a store to T storage keeps
  return V3{(double)(T)a.x, (double)(T)a.y, (double)(T)a.z};
}

// computeGradients(g) + proportional (MPRGPSolver.h:163-181) for one row at
// gradient g (as stored), face f and iterate x: phi and beta stored, the
// loop head's sums in slots 0-2.
template <class T>
__device__ __forceinline__ void loop_head_row(const Vecs<T> &v, const Row &R, bool f, V3 g, V3 x,
                                              double alpha_bar, int r, double (&P)[kNumPart]) {
  const V3 gs = round_to<T>(g);
  const V3 ph = phi_of(R, f, gs);
  const V3 be = beta_of(R, f, gs, ph);
  st3(v.phi, r, ph);
  st3(v.beta, r, be);
  const V3 gp = add(ph, be);
  P[0] += dot3(gp, gp);
  P[1] += dot3(be, be);
  const V3 px = proj(R, axmy(x, alpha_bar, ph));
  P[2] += dot3(sub(x, px), ph);
}

// The column indices the SpMVs read: 16-bit (C16, matrices of at most
// 65536 nodes) or 32-bit; the same columns either way.
template <bool C16>
__device__ __forceinline__ auto cols(const QpDev &q) {
  if constexpr (C16)
    return q.col16;
  else
    return q.col;
}

template <class T, bool MX, int UB, bool NT, bool C16>
__device__ __forceinline__ void row_work(const QpDev &q, const Vecs<T> &vin, int s, int lane, int r, int op,
                                         const QpState &st, double (&P)[kNumPart]) {
  const bool mixed = MX && st.mixed;
  // p and p2, beta and beta2 swap roles with st.pb, st.bb
  Vecs<T> v = vin;
  if (st.pb) {
    v.p = vin.p2;
    v.p2 = vin.p;
  }
  if (st.bb) {
    v.beta = vin.beta2;
    v.beta2 = vin.beta;
  }
  // x and y swap buffers when a CG step is accepted (x = y without a copy):
  // st.xb names the buffer that holds x
  double *const xc = st.xb ? v.y : v.x;
  double *const yc = st.xb ? v.x : v.y;
  const bool pre = st.precond && st.has_rows;
  switch (op) {
    case OP_INIT:
    case OP_EXP3:
    case OP_CL_EXP2: {
      // g = A x - b ; DECIDE_FACE(x) ; phi ; z ; p = z   (MPRGPSolver.h:148-156, 263-268, 210-215)
      // The mixed form's INIT starts a correction QP from d = 0, which
      // R_START wrote: A d is then +0 in every entry (finite values times
      // zero, summed from +0), so the SpMV is skipped with the same bits.
      V3 acc[1] = {V3{0.0, 0.0, 0.0}};
      const double *x = xc;
      if (!(mixed && op == OP_INIT))
        spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(x, j); },
                            acc);
      V3 g = sub(acc[0], ld3(v.b, r));
      const V3 xr = ld3(xc, r);  // in flight with the row flag
      Row R = ldrow(q, v, r);
      bool f = face_of(R, xr);
      V3 ph = phi_of(R, f, g);
      V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(v.g, r, g);
      st3(v.z, r, z);
      st3(v.p, r, z);
      gmem(q.face)[r] = f;
      if (op == OP_EXP3) st3(yc, r, ld3(v.py, r));
      // INIT and EXP3 end with the next outer iteration's head (its phi
      // replaces this one); the classic CL_EXP2 is followed by TOP
      if (op == OP_CL_EXP2) st3(v.phi, r, ph);
      else loop_head_row(v, R, f, g, xr, st.alpha_bar, r, P);
      break;
    }
    case OP_TOP: {
      // computeGradients (MPRGPSolver.h:163-173) + proportional (:174-181)
      Row R = ldrow(q, v, r);
      bool f = gmem(q.face)[r];
      if (st.p_pending) {  // classic CGStep tail: p = z - beta p (MPRGPSolver.h:192-193)
        st3(v.p, r, axmy(ld3(v.z, r), st.beta_cg, ld3(v.p, r)));
      }
      loop_head_row(v, R, f, ld3(v.g, r), ld3(xc, r), st.alpha_bar, r, P);
      break;
    }
    case OP_CG_AP:
    case OP_CL_AP: {
      V3 acc[1];
      const T *p = v.p;
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(p, j); }, acc);
      st3(v.ap, r, acc[0]);
      V3 pr = ld3(v.p, r);
      P[0] += dot3(pr, acc[0]);
      const V3 gr = ld3(v.g, r);
      P[1] += dot3(ld3(v.z, r), gr);
      if (mixed && st.cg_entry) P[2] += dot3(ld3(xc, r), sub(gr, ld3(v.b, r)));
      if (op == OP_CL_AP) {
        // stepLimit (MPRGPProjection.h:263-282)
        Row R = ldrow(q, v, r);
        double lim = kDblMax;
        if (R.has) {
          V3 xr = ld3(xc, r);
          double jx = dot3(R.n, xr), jd = dot3(R.n, pr);
          if (jd > kEps && jx > R.c) lim = (jx - R.c) / jd;
        }
        P[kMinSlot] = fmin(P[kMinSlot], lim);
      }
      break;
    }
    case OP_CG_PY: {
      const V3 xr = ld3(xc, r), pr = ld3(v.p, r);  // in flight with the row flag
      Row R = ldrow(q, v, r);
      V3 y = axmy(xr, st.alpha, pr);
      st3(yc, r, y);
      st3(v.py, r, proj(R, y));
      P[0] += infeasible(R, y);
      break;
    }
    case OP_CG_PYG: {
      // Mixed inner solve, CGMonotonicStep (MPRGPSolver.h:281-297) in one
      // pass: the trial point with the sums for f(y), and the gradient
      // update of an accepted step applied in place.  A rejected step leaves
      // the CG loop through EXP1/EXP2, PROP1 or FINAL, and the first two
      // recompute g = A x - b, phi and z before reading them (the reference
      // does the same: MPRGPSolver.h:255-261, 435-437); FINAL reads x only.
      // The face stays fixed inside the CG loop.
      if (!MX) break;
      // the row's own x and p first, so their loads are in flight while the
      // row flag (and, on rows, the row data behind it) arrives
      const bool f = gmem(q.face)[r];
      const V3 xr = ld3(xc, r), pr = ld3(v.p, r);
      Row R = ldrow(q, v, r);
      const V3 y = axmy(xr, st.alpha, pr);
      st3(yc, r, y);
      // f(P(y)) = f(y) + d'(A y - b) + d'A d/2 for d = P(y) - y
      // (projectFunValue, MPRGPSolver.h:245-249), with d taken on the rows
      // that violate their plane by more than the feasibility tolerance
      // (isFeasible's eps): the rows on their plane, whose round-off P would
      // move by ~1e-17, count as feasible.  A violated row gathers every
      // neighbour's d by the same rule, formed from its x, p and row (none of
      // which this launch writes, all loaded at once); no other row reads the
      // matrix.
      const bool viol = infeasible(R, y) != 0.0;
      P[0] += viol ? 1.0 : 0.0;
      V3 dr{0.0, 0.0, 0.0};
      double dad = 0.0;
      if (viol) {
        const double lam = (R.c - dot3(R.n, y)) / R.jj;
        dr = V3{R.n.x * lam, R.n.y * lam, R.n.z * lam};
        const double *xx = xc;
        const T *pp = v.p;
        const double alpha = st.alpha;
        V3 acc[1];
        spmv_row<1, 2, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) {
          const V3 yj = axmy(ld3(xx, j), alpha, ld3(pp, j));
          const V3 nj = ld3(v.cn, j);
          const double vj = gmem(v.cc)[j] - dot3(nj, yj);
          const bool on = gmem(q.cf)[j] != 0 && vj > kEps;
          const double lj = vj / gmem(v.cj)[j];
          xs[0] = on ? V3{nj.x * lj, nj.y * lj, nj.z * lj} : V3{0.0, 0.0, 0.0};
        }, acc);
        dad = dot3(dr, acc[0]);
      }
      const V3 gr = ld3(v.g, r), ap = ld3(v.ap, r);
      P[1] += dot3(pr, gr);
      const V3 g = axmy(gr, st.alpha, ap);
      const V3 ph = phi_of(R, f, g);
      const V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(v.g, r, g);
      st3(v.phi, r, ph);
      st3(v.z, r, z);
      // sums over the values later reads of z, g and phi see
      const V3 gs = round_to<T>(g), phs = round_to<T>(ph);
      P[2] += dot3(round_to<T>(z), gs);
      P[3] += dot3(z, ap);
      // computeGradients(g, false) and proportional at x = y; beta goes to
      // the other buffer, which only an accepted step makes current (a
      // rejected step's proportioning reads the old beta)
      const V3 be = beta_of(R, f, gs, phs);
      st3(v.beta2, r, be);
      const V3 gp = add(phs, be);
      P[4] += dot3(gp, gp);
      P[5] += dot3(be, be);
      const V3 px = proj(R, axmy(y, st.alpha_bar, phs));
      P[6] += dot3(sub(y, px), phs);
      P[7] += dot3(dr, gs) + 0.5 * dad;  // A y - b = g - alpha AP, as stored
      break;
    }
    case OP_CG_BAP: {
      // p = z - beta p and the next step's A p (MPRGPSolver.h:298, 304-306)
      // in one pass.  p' is formed again at every gather from z and p, which
      // no tile writes in this launch (p' goes to the other p buffer),
      // rounded as its store.
      const T *zz = v.z, *pp = v.p;
      const double beta = st.beta_cg;
      V3 acc[1];
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) {
        xs[0] = round_to<T>(axmy(ld3(zz, j), beta, ld3(pp, j)));
      }, acc);
      const V3 pn = round_to<T>(axmy(ld3(v.z, r), beta, ld3(v.p, r)));
      st3(v.p2, r, pn);
      st3(v.ap, r, acc[0]);
      P[0] += dot3(pn, acc[0]);
      break;
    }
    case OP_CG_Y: {
      const double *py = v.py;
      V3 pyr = ld3(v.py, r);
      V3 br = ld3(v.b, r);
      if (st.cg_entry) {
        const double *x = xc;
        V3 acc[2];
        spmv_row<2, 1, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[2]) {
          xs[0] = ld3(py, j);
          xs[1] = ld3(x, j);
        }, acc);
        V3 xr = ld3(xc, r);
        P[0] += dot3(pyr, acc[0]);
        P[1] += dot3(pyr, br);
        P[2] += dot3(xr, acc[1]);
        P[3] += dot3(xr, br);
      } else {
        V3 acc[1];
        spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(py, j); }, acc);
        P[0] += dot3(pyr, acc[0]);
        P[1] += dot3(pyr, br);
      }
      break;
    }
    case OP_CG_YA: {
      // CGMonotonicStep (MPRGPSolver.h:283-302) after the CG entry: f(P(y))
      // by A P(y) as CG_Y, then CG_A's update applied speculatively at x = y.
      // A rejected step leaves the CG loop through EXP1/EXP2, PROP1 or FINAL,
      // and the first two recompute g, phi and z before reading them
      // (MPRGPSolver.h:255-261, 435-437); FINAL reads x only; beta goes to
      // the other buffer, current only once the step is accepted.
      const double *py = v.py;
      V3 acc[1];
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(py, j); }, acc);
      const V3 pyr = ld3(v.py, r);
      P[0] += dot3(pyr, acc[0]);
      P[1] += dot3(pyr, ld3(v.b, r));
      const bool f = gmem(q.face)[r];  // in flight with the row flag
      const V3 ap = ld3(v.ap, r), gr = ld3(v.g, r);
      Row R = ldrow(q, v, r);
      const V3 g = axmy(gr, st.alpha, ap);
      const V3 ph = phi_of(R, f, g);
      const V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(v.g, r, g);
      st3(v.phi, r, ph);
      st3(v.z, r, z);
      P[2] += dot3(z, ap);
      P[3] += dot3(z, g);
      const V3 be = beta_of(R, f, g, ph);
      st3(v.beta2, r, be);
      const V3 gp = add(ph, be);
      P[4] += dot3(gp, gp);
      P[5] += dot3(be, be);
      const V3 y = ld3(yc, r);
      const V3 px = proj(R, axmy(y, st.alpha_bar, ph));
      P[6] += dot3(sub(y, px), ph);
      break;
    }
    case OP_CG_A: {
      // CGMonotonicStep body (MPRGPSolver.h:293-302) but p = z - beta p,
      // which needs beta first and goes with the next A p (CG_BAP)
      bool f = gmem(q.face)[r];  // in flight with the row flag
      V3 ap = ld3(v.ap, r);      // x = y happened in the control (buffer swap)
      const V3 gr = ld3(v.g, r);
      Row R = ldrow(q, v, r);
      V3 g = axmy(gr, st.alpha, ap);
      V3 ph = phi_of(R, f, g);
      V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(v.g, r, g);
      st3(v.phi, r, ph);
      st3(v.z, r, z);
      P[0] += dot3(z, ap);
      P[1] += dot3(z, g);  // the next step's alpha numerator
      // computeGradients(g, false) and proportional at the new x
      const V3 be = beta_of(R, f, g, ph);
      st3(v.beta, r, be);
      const V3 gp = add(ph, be);
      P[2] += dot3(gp, gp);
      P[3] += dot3(be, be);
      const V3 x = ld3(xc, r);
      const V3 px = proj(R, axmy(x, st.alpha_bar, ph));
      P[4] += dot3(sub(x, px), ph);
      break;
    }
    case OP_EXP1: {
      Row R = ldrow(q, v, r);
      st3(yc, r, proj(R, ld3(xc, r)));
      break;
    }
    case OP_EXP2: {
      // ExpMonotonicStep, first half (MPRGPSolver.h:255-261): the new y goes
      // to py because neighbours still gather y in this launch.
      V3 acc[1];
      const double *y = yc;
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(y, j); }, acc);
      V3 g = sub(acc[0], ld3(v.b, r));
      Row R = ldrow(q, v, r);
      V3 yr = ld3(yc, r);
      bool f = face_of(R, yr);
      V3 ph = phi_of(R, f, g);
      V3 yn = axmy(yr, st.alpha_bar, ph);
      st3(v.g, r, g);
      st3(v.py, r, yn);
      st3(xc, r, proj(R, yn));
      gmem(q.face)[r] = f;
      P[0] += infeasible(R, yn);
      break;
    }
    case OP_PROP1: {
      // g = A x - b ; AD = A beta ; beta.AD ; g.beta  (MPRGPSolver.h:435-437, 224-231)
      V3 acc[2];
      const double *x = xc;
      const T *be = v.beta;
      spmv_row<2, 1, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[2]) {
        xs[0] = ld3(x, j);
        xs[1] = ld3(be, j);
      }, acc);
      V3 g = sub(acc[0], ld3(v.b, r));
      st3(v.g, r, g);
      st3(v.ap, r, acc[1]);
      V3 d = ld3(v.beta, r);
      P[0] += dot3(d, acc[1]);
      P[1] += dot3(g, d);
      break;
    }
    case OP_CL_PROP1: {
      V3 acc[1];
      const T *be = v.beta;
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(be, j); }, acc);
      st3(v.ap, r, acc[0]);
      V3 d = ld3(v.beta, r);
      P[0] += dot3(d, acc[0]);
      P[1] += dot3(ld3(v.g, r), d);
      break;
    }
    case OP_PROP2: {
      // PropStep tail (MPRGPSolver.h:232-241)
      Row R = ldrow(q, v, r);
      V3 xn = proj(R, axmy(ld3(xc, r), st.alpha, ld3(v.beta, r)));
      V3 g = axmy(ld3(v.g, r), st.alpha, ld3(v.ap, r));
      bool f = face_of(R, xn);
      V3 ph = phi_of(R, f, g);
      V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(xc, r, xn);
      st3(v.g, r, g);
      st3(v.z, r, z);
      st3(v.p, r, z);
      gmem(q.face)[r] = f;
      loop_head_row(v, R, f, g, xn, st.alpha_bar, r, P);  // the next iteration's head
      break;
    }
    case OP_CL_CG: {
      // CGStep (MPRGPSolver.h:188-191); p is updated at the next TOP
      Row R = ldrow(q, v, r);
      bool f = gmem(q.face)[r];
      V3 ap = ld3(v.ap, r);
      st3(xc, r, axmy(ld3(xc, r), st.alpha, ld3(v.p, r)));
      V3 g = axmy(ld3(v.g, r), st.alpha, ap);
      V3 ph = phi_of(R, f, g);
      V3 z = precond_of(v, pre, R, f, r, g, ph);
      st3(v.g, r, g);
      st3(v.phi, r, ph);
      st3(v.z, r, z);
      P[0] += dot3(z, ap);
      break;
    }
    case OP_CL_EXP: {
      // ExpStep first half (MPRGPSolver.h:202-208)
      Row R = ldrow(q, v, r);
      V3 xt = axmy(ld3(xc, r), st.alpha_f, ld3(v.p, r));
      V3 g = axmy(ld3(v.g, r), st.alpha_f, ld3(v.ap, r));
      bool f = face_of(R, xt);
      V3 ph = phi_of(R, f, g);
      xt = axmy(xt, st.alpha_bar, ph);
      st3(xc, r, proj(R, xt));
      st3(v.g, r, g);
      gmem(q.face)[r] = f;
      break;
    }
    case OP_FINAL: {
      // the result goes to the first buffer, where the caller reads x
      V3 xr = ld3(xc, r);
      if (st.variant != VAR_CLASSIC) xr = proj(ldrow(q, v, r), xr);
      st3(v.x, r, xr);
      P[0] += gmem(q.face)[r] ? 1.0 : 0.0;
      break;
    }
    default:
      break;
  }
}

// Ops on FP64 vectors in both forms: projection of the warm start, the power
// method (its SpMV reads the matrix of the QP's form), and the outer
// refinement of the mixed form.
template <bool MX, int UB, bool NT, bool C16>
__device__ __forceinline__ void row_work_fp64(const QpDev &q, int s, int lane, int r, int op,
                                              const QpState &st, double (&P)[kNumPart]) {
  const Vecs<double> &v = q.d;
  switch (op) {
    case OP_PROJX: {
      const V3 xr = ld3(v.x, r);  // in flight with the row flag
      Row R = ldrow(q, v, r);
      st3(v.x, r, proj(R, xr));
      break;
    }
    case OP_POW_IT: {
      // Iteration k: v_k = w_{k-1}/nu_{k-1} (v_0 is the start vector);
      // |v_k - v_{k-1}|^2 closes iteration k-1; w_k = A v_k, formed as
      // (A w_{k-1})/nu_{k-1} so that the gathers read w_{k-1}, which no tile
      // writes in this launch (equal to A v_k up to rounding).  v and w each
      // alternate between two buffers with the parity of k.  The mixed form
      // multiplies by the FP32 copy of A, the matrix its inner solve
      // iterates with (vectors and norms stay FP64).
      const int k = st.power_it;
      double *vk = (k & 1) ? q.pw : q.pv;
      const double *vp = (k & 1) ? q.pv : q.pw;
      double *wk = (k & 1) ? q.pz : q.pu;
      const double *wp = (k & 1) ? q.pu : q.pz;
      const double nu = st.nu;
      if (k >= 1) {
        const V3 a = ld3(wp, r);
        const V3 vr = V3{a.x / nu, a.y / nu, a.z / nu};
        st3(vk, r, vr);
        const V3 dlt = sub(vr, ld3(vp, r));
        P[1] += dot3(dlt, dlt);
      }
      if (k > 1000) break;  // the last launch only closes iteration 1000
      const double *in = k == 0 ? vk : wp;
      V3 acc[1];
      auto gather = [&](int j, V3 (&xs)[1]) { xs[0] = ld3(in, j); };
      if (MX && st.mixed)
        spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), q.f.val, s, lane, gather, acc);
      else
        spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, gather, acc);
      const V3 w = k == 0 ? acc[0] : V3{acc[0].x / nu, acc[0].y / nu, acc[0].z / nu};
      st3(wk, r, w);
      P[0] += dot3(w, w);
      break;
    }
    case OP_POW_SAVE: {
      const double *pv = (st.power_it & 1) ? q.pw : q.pv;
      st3(q.ev, r, ld3(pv, r));
      break;
    }
    case OP_R_START: {
      // FP64 residual and KKT norm at x; the correction QP
      //   min 1/2 d'Ad - r'd  s.t.  n.d >= c - n.x,  r = b - A x
      // is set up in the FP32 storage with d = 0.
      V3 acc[1];
      const double *x = v.x;
      spmv_row<1, UB, NT>(q.sl_off, cols<C16>(q), v.val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(x, j); }, acc);
      V3 g = sub(acc[0], ld3(v.b, r));
      V3 xr = ld3(v.x, r);  // in flight with the row flag
      Row R = ldrow(q, v, r);
      bool f = face_of(R, xr);
      V3 ph = phi_of(R, f, g);
      V3 be = beta_of(R, f, g, ph);
      V3 gp = add(ph, be);
      P[0] += dot3(gp, gp);
      gmem(q.face)[r] = f;
      st3(q.f.b, r, V3{-g.x, -g.y, -g.z});
      st3(q.f.x, r, V3{0.0, 0.0, 0.0});
      if (R.has) gmem(q.f.cc)[r] = R.c - dot3(R.n, xr);
      break;
    }
    case OP_R_UPD: {
      st3(v.x, r, add(ld3(v.x, r), ld3(q.f.x, r)));
      break;
    }
    case OP_R_FINAL: {
      const V3 xr = ld3(v.x, r);  // in flight with the row flag
      P[0] += gmem(q.face)[r] ? 1.0 : 0.0;
      Row R = ldrow(q, v, r);
      st3(v.x, r, proj(R, xr));
      break;
    }
    default:
      break;
  }
}

__device__ __forceinline__ bool fp64_op(int op) {
  return op == OP_PROJX || op == OP_POW_IT || op == OP_POW_SAVE ||
         op == OP_R_START || op == OP_R_UPD || op == OP_R_FINAL;
}

// One slice of a tile per wavefront.
template <int TS, bool MX, int UB, bool NT, bool C16>
__device__ __forceinline__ void slice_work(const QpDev &q, int s, int lane, int op, const QpState &st,
                                           double (&P)[kNumPart]) {
  const int r = s * kSlice + lane;
  if (r >= q.n_nodes) return;
  if (fp64_op(op)) {
    row_work_fp64<MX, UB, NT, C16>(q, s, lane, r, op, st, P);
  } else if (MX && st.mixed) {
    row_work<float, MX, UB, NT, C16>(q, q.f, s, lane, r, op, st, P);
  } else {
    row_work<double, MX, UB, NT, C16>(q, q.d, s, lane, r, op, st, P);
  }
}

// A tile is up to TS * SPW consecutive slices: wave w streams the tile's
// slices w, w + TS, ..., w + (SPW - 1) TS one after the other (SPW slices
// per wavefront, the call's slices_per_wave), adding each row's sums to its
// lane's partials in that order.  SPW is 1 or 2; the second slice is a
// second inlined call rather than a loop, whose body the register allocator
// spilled (110 VGPRs against 3).
template <int TS, bool MX, int UB, bool NT, int SPW, bool C16>
__device__ void tile_work(const QpDev &q, const Tile &t, int op, const QpState &st, double (&P)[kNumPart]) {
  const int wave = threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  if constexpr (SPW == 1) {
    if (wave < t.slices) slice_work<TS, MX, UB, NT, C16>(q, t.slice0 + wave, lane, op, st, P);
  } else {
    static_assert(SPW == 2, "one or two slices per wavefront");
    if (wave < t.slices) slice_work<TS, MX, UB, NT, C16>(q, t.slice0 + wave, lane, op, st, P);
    if (wave + TS < t.slices) slice_work<TS, MX, UB, NT, C16>(q, t.slice0 + wave + TS, lane, op, st, P);
  }
}

// Partial sums of one QP's tiles, slots J0..J0+kGroup-1, summed per lane:
// every load of a lane's first two columns is issued before the first use.
constexpr int kGroup = 4;
template <int J0>
__device__ __forceinline__ void sum_part(const double *pp, int nt, int lane, double (&acc)[kNumPart],
                                         double &mn) {
  double a[2][kGroup];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int k = lane + u * kSlice;
#pragma unroll
    for (int j = 0; j < kGroup; ++j) a[u][j] = (k < nt) ? gmem(pp)[(size_t)(J0 + j) * nt + k] : 0.0;
  }
#pragma unroll
  for (int j = 0; j < kGroup; ++j) acc[J0 + j] = a[0][j] + a[1][j];
  if (J0 == 0) {
#pragma unroll
    for (int j = kGroup; j < kNumPart; ++j) acc[j] = 0.0;
    mn = fmin(lane < nt ? a[0][kMinSlot] : kDblMax, lane + kSlice < nt ? a[1][kMinSlot] : kDblMax);
  }
  // QPs of more than 128 tiles: rounds of four columns per lane, every load
  // of a round issued before its sums, then one last round of the columns
  // left.  Which columns a round holds is the same on every lane (a column of
  // 64 tiles exists if its first tile does); in the last round lanes past the
  // last tile load that tile and do not add it.  Each lane adds its columns
  // in order, as one column at a time would.
  int k0 = 2 * kSlice;
  for (; k0 + 4 * kSlice <= nt; k0 += 4 * kSlice) {
    double v[4][kGroup];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < kGroup; ++j) v[u][j] = gmem(pp)[(size_t)(J0 + j) * nt + k0 + u * kSlice + lane];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int j = 0; j < kGroup; ++j) {
        acc[J0 + j] += v[u][j];
        if (J0 == 0 && j == kMinSlot) mn = fmin(mn, v[u][j]);
      }
  }
  if (k0 < nt) {
    double v[4][kGroup];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u * kSlice >= nt) break;
      const int k = min(k0 + u * kSlice + lane, nt - 1);
#pragma unroll
      for (int j = 0; j < kGroup; ++j) v[u][j] = gmem(pp)[(size_t)(J0 + j) * nt + k];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (k0 + u * kSlice >= nt) break;
      const bool in = k0 + u * kSlice + lane < nt;
#pragma unroll
      for (int j = 0; j < kGroup; ++j) {
        acc[J0 + j] = in ? acc[J0 + j] + v[u][j] : acc[J0 + j];
        if (J0 == 0 && j == kMinSlot) mn = in ? fmin(mn, v[u][j]) : mn;
      }
    }
  }
}

// A QpState copy moves as 64 words of 8 bytes, one per lane; the type may
// alias the state's int and double fields.
typedef uint64_t __attribute__((__may_alias__)) StateWord;

#ifndef MPRGP_STEP_MIN_WAVES
#define MPRGP_STEP_MIN_WAVES 4
#endif

// The one-vector SpMVs' unroll: the blocks of UB rows' slots are loaded per
// round.  The sums still run over the blocks in order, so UB changes no bit
// of a result.  kUnrollTail, with more registers and fewer waves, is for
// launches with few tiles, where a lone row's chain of load rounds sets the
// launch time.
#ifndef MPRGP_STEP_UNROLL
#define MPRGP_STEP_UNROLL 3
#endif
constexpr int kUnroll = MPRGP_STEP_UNROLL;
constexpr int kUnrollTail = 6;

// MX: the batch holds QPs of the mixed form (their inner ops run on the FP32
// copies); FP64 QPs of the same batch take the FP64 path.  C16: every QP of
// the batch has 16-bit column indices (QpDev::col16).
template <int TS, bool MX, int UB, bool NT, int SPW, bool C16>
__global__ void __launch_bounds__(TS * kSlice, UB == kUnroll ? MPRGP_STEP_MIN_WAVES : 2)
    mprgp_step(const QpDev *__restrict__ qps,
                                                           const Tile *__restrict__ tiles,
                                                           int *running, int parity) {
  __shared__ double red[TS][kNumPart];
  __shared__ QpState sst;
  const int wave = threadIdx.x / kSlice;
  const int lane = threadIdx.x % kSlice;
  // one tile per workgroup: the table has one slot per workgroup, laid out
  // for the XCDs, and the workgroups of unused slots exit
  const Tile t = tiles[blockIdx.x];
  if (t.qp < 0) return;
  const QpDev &q = qps[t.qp];
  // 1. state the previous launch executed + its partials -> this launch's op.
  //    Wave 0 issues every load (one 8-byte state word per lane, the
  //    partials slot-major: [parity][slot][tile]) before the first wait.
  if (wave == 0) {
    const StateWord w = gmem(reinterpret_cast<const StateWord *>(t.st + (parity ^ 1)))[lane];
    const int nt = t.nt;
    const double *pp = t.part + (size_t)(parity ^ 1) * kNumPart * nt;
    double acc[kNumPart];
    double mn;
    sum_part<0>(pp, nt, lane, acc, mn);
    reinterpret_cast<StateWord *>(&sst)[lane] = w;
    // the op whose partials these are: the state's first word, held by lane 0
    const int prev = __builtin_amdgcn_readfirstlane((int)(uint32_t)w);
    const int ns = sum_slots(prev);
    if (ns > kGroup) sum_part<kGroup>(pp, nt, lane, acc, mn);  // the wide ops' upper slots
    Sums S;
#pragma unroll
    for (int j = 0; j < kNumPart; ++j) S.v[j] = (j < ns) ? wave_sum(acc[j]) : 0.0;
    if (prev == OP_CL_AP) S.v[kMinSlot] = wave_min(mn);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      const bool skip = sst.fresh || sst.op == OP_DONE;
      const bool was_done = sst.op == OP_DONE;
      if (!skip) control(sst, S);
      sst.fresh = 0;
      if (t.local == 0) {  // the QP's first tile counts the launch (or its end)
        if (sst.op != OP_DONE) {
          sst.steps++;
          sst.opcnt[sst.op]++;
        } else if (!was_done) {
          atomicSub(running, 1);
          *gmem(q.done_flag) = q.tag;
        }
      }
    }
  }
  __syncthreads();
  const QpState &st = sst;
  const int op = st.op;
  if (op != OP_DONE) {
    const bool has_min = op == OP_CL_AP;
    double P[kNumPart];
#pragma unroll
    for (int j = 0; j < kNumPart; ++j) P[j] = 0.0;
    if (has_min) P[kMinSlot] = kDblMax;
    tile_work<TS, MX, UB, NT, SPW, C16>(q, t, op, st, P);
    // wave then block reduction in a fixed order, over the op's slots only
    const int ns = has_min ? kMinSlot + 1 : sum_slots(op);
    if (ns > 0) {
#pragma unroll
      for (int j = 0; j < kNumPart; ++j)
        if (j < ns) P[j] = (has_min && j == kMinSlot) ? wave_min(P[j]) : wave_sum(P[j]);
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < kNumPart; ++j) red[wave][j] = P[j];
      }
      __syncthreads();
      if (wave == 0 && lane < ns) {
        const int j = lane;
        const bool m = has_min && j == kMinSlot;
        double acc = red[0][j];
        for (int w = 1; w < TS; ++w) acc = m ? fmin(acc, red[w][j]) : acc + red[w][j];
        gmem(t.part)[((size_t)parity * kNumPart + j) * t.nt + t.local] = acc;
      }
    }
  }
  // 2. the QP's first tile persists the state this launch executes (wave 0
  //    wrote it in the prologue; the other waves only read it)
  if (t.local == 0 && wave == 0)
    gmem(reinterpret_cast<StateWord *>(t.st + parity))[lane] = reinterpret_cast<const StateWord *>(&sst)[lane];
}

// ------------------------------------------------------- helper kernels

// out[i] = in[perm[i]]: user node order -> internal order.
__global__ void gather_nodes(const double *__restrict__ in, double *__restrict__ out,
                             const int *__restrict__ perm, int n_nodes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  int o = perm[i];
  out[3 * (size_t)i + 0] = in[3 * (size_t)o + 0];
  out[3 * (size_t)i + 1] = in[3 * (size_t)o + 1];
  out[3 * (size_t)i + 2] = in[3 * (size_t)o + 2];
}
// out[perm[i]] = in[i]: internal order -> user node order.
__global__ void scatter_nodes(const double *__restrict__ in, double *__restrict__ out,
                              const int *__restrict__ perm, int n_nodes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  int o = perm[i];
  out[3 * (size_t)o + 0] = in[3 * (size_t)i + 0];
  out[3 * (size_t)o + 1] = in[3 * (size_t)i + 1];
  out[3 * (size_t)o + 2] = in[3 * (size_t)i + 2];
}

// dinv = 1/diag(A) from the diagonal block of each row (MPRGPPrecondition.h:44-49).
__global__ void setup_diag(const int *__restrict__ sl_off, const int *__restrict__ col,
                           const double *__restrict__ val, double *__restrict__ dinv, int n_nodes) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_nodes) return;
  int s = r / kSlice, lane = r % kSlice;
  int b0 = sl_off[s];
  int W = (sl_off[s + 1] - b0) / kSlice;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int k = 0; k < W; ++k) {
    if (col[b0 + k * kSlice + lane] == r) {
      const double *v = val + (size_t)b0 * 9 + lane + (size_t)k * 9 * kSlice;
      d0 = v[0 * kSlice];
      d1 = v[4 * kSlice];
      d2 = v[8 * kSlice];
      break;
    }
  }
  dinv[3 * (size_t)r + 0] = 1.0 / d0;
  dinv[3 * (size_t)r + 1] = 1.0 / d1;
  dinv[3 * (size_t)r + 2] = 1.0 / d2;
}

// diag(A) from the diagonal block of each row (getDiagonal, MPRGPUtility.h:503-515).
__global__ void get_diag(const int *__restrict__ sl_off, const int *__restrict__ col,
                         const double *__restrict__ val, double *__restrict__ diag, int n_nodes) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_nodes) return;
  int s = r / kSlice, lane = r % kSlice;
  int b0 = sl_off[s];
  int W = (sl_off[s + 1] - b0) / kSlice;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int k = 0; k < W; ++k) {
    if (col[b0 + k * kSlice + lane] == r) {
      const double *v = val + (size_t)b0 * 9 + lane + (size_t)k * 9 * kSlice;
      d0 = v[0 * kSlice];
      d1 = v[4 * kSlice];
      d2 = v[8 * kSlice];
      break;
    }
  }
  diag[3 * (size_t)r + 0] = d0;
  diag[3 * (size_t)r + 1] = d1;
  diag[3 * (size_t)r + 2] = d2;
}

// One partial of a.b per workgroup, for reduce_parts.
__global__ void __launch_bounds__(256) dot_parts(const double *__restrict__ a, const double *__restrict__ b,
                                                 size_t n, double *__restrict__ part) {
  __shared__ double red[256];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  red[threadIdx.x] = i < n ? a[i] * b[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Several dot products at once: pair blockIdx.y, its partials at
// part[blockIdx.y * gridDim.x ...], each formed as dot_parts forms it.
struct DotPairs {
  const double *a[kMaxDots], *b[kMaxDots];
};
__global__ void __launch_bounds__(256) dot_parts_n(DotPairs pr, size_t n, double *__restrict__ part) {
  __shared__ double red[256];
  const int j = blockIdx.y;
  const double *__restrict__ a = pr.a[j];
  const double *__restrict__ b = pr.b[j];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  red[threadIdx.x] = i < n ? a[i] * b[i] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)j * gridDim.x + blockIdx.x] = red[0];
}

// cs = 1/((n o dinv).n) per row (MPRGPPrecondition.h:81-88).
__device__ __forceinline__ double row_factor(const double *dinv, const double *cn, unsigned char cf, int r) {
  if (!cf) return 0.0;
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    double n = cn[3 * (size_t)r + d];
    s += (n * dinv[3 * (size_t)r + d]) * n;
  }
  return 1.0 / s;
}

__global__ void setup_rows(const double *__restrict__ dinv, const double *__restrict__ cn,
                           const unsigned char *__restrict__ cf, double *__restrict__ cs, int n_nodes) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_nodes) return;
  cs[r] = row_factor(dinv, cn, cf[r], r);
}

// setup_rows for every QP of a batch table in one launch: blockIdx.y is the
// QP, blockIdx.x a run of 256 of its nodes.  cs is the QP's own factor array
// (read-only to the step kernel, written here).  A QP whose descriptor names
// an x0 also restarts from it: x = x0 on the same nodes.
__global__ void __launch_bounds__(256) setup_rows_batch(const QpDev *__restrict__ qps) {
  const QpDev &q = qps[blockIdx.y];
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= q.n_nodes) return;
  const_cast<double *>(q.d.cs)[r] = row_factor(q.d.dinv, q.d.cn, q.cf[r], r);
  if (q.x0) {
    const size_t i = 3 * (size_t)r;
    q.d.x[i] = q.x0[i];
    q.d.x[i + 1] = q.x0[i + 1];
    q.d.x[i + 2] = q.x0[i + 2];
  }
}

// Decoupled plane collision at the current iterate: node r gets the row of the
// first plane k (in the given order) with n_k.x_r + p_k - delta_k <= 0, i.e.
// J row = n_k on the node's 3 columns and c = -p_k (LinearConCollider.cpp:140-161
// per plane, MprgpFemSolver.cpp:268-271 over the planes, getConstraints
// LinearConCollider.cpp:57-72); every other node's row storage is cleared.  With
// one row per node, the reference's per-node list of distinct planes (the 1e-4
// test of addConPlane) never changes which plane wins.  planes: 5 doubles each
// {nx, ny, nz, p, delta}.  count must be zero on entry; it receives the rows.
__global__ void __launch_bounds__(256) collide_planes(const double *__restrict__ x,
                                                      const double *__restrict__ planes, int n_planes,
                                                      double *__restrict__ cn, double *__restrict__ cc,
                                                      double *__restrict__ cj, unsigned char *__restrict__ cf,
                                                      int *__restrict__ count, int n_nodes) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (r < n_nodes) {
    const V3 xr = ld3(x, r);
    V3 n = {0.0, 0.0, 0.0};
    double c = 0.0;
    for (int k = 0; k < n_planes; ++k) {
      const double *pl = planes + 5 * k;
      // n.dot(pos) + plane[3] - delta, the dot summed left to right
      double d = pl[0] * xr.x + pl[1] * xr.y;
      d = d + pl[2] * xr.z;
      d = d + pl[3];
      d = d - pl[4];
      if (d <= 0.0) {
        hit = true;
        n = V3{pl[0], pl[1], pl[2]};
        c = -pl[3];
        break;
      }
    }
    st3(cn, r, n);
    cc[r] = c;
    cj[r] = hit ? (n.x * n.x + n.y * n.y) + n.z * n.z : 0.0;
    cf[r] = hit ? 1 : 0;
  }
  const unsigned long long b = __ballot(hit);
  if ((threadIdx.x % kSlice) == 0 && b != 0) atomicAdd(count, (int)__popcll(b));
}

// The projector and preconditioner of a QP's rows applied once, per node,
// with the step kernel's own per-node functions (proj, infeasible, face_of,
// phi_of, beta_of, precond_of and CL_AP's step-limit rule).  Vectors are in
// internal node order.  Ops that reduce write one partial per workgroup into
// part (a sum, or a minimum for PJ_STEP_LIMIT), which reduce_parts finishes in
// a fixed order.
__global__ void __launch_bounds__(256) projector_op(ProjArgs g) {
  __shared__ double red[256];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const bool is_min = g.op == PJ_STEP_LIMIT;
  double acc = is_min ? kDblMax : 0.0;
  if (r < g.n_nodes) {
    Row R;
    R.has = gmem(g.cf)[r] != 0;
    R.n = R.has ? ld3(g.cn, r) : V3{0.0, 0.0, 0.0};
    R.c = R.has ? gmem(g.cc)[r] : 0.0;
    R.jj = R.has ? gmem(g.cj)[r] : 0.0;
    R.s = R.has ? gmem(g.cs)[r] : 0.0;
    const V3 a = ld3(g.a, r);
    switch (g.op) {
      case PJ_PROJECT: st3(g.out, r, proj(R, a)); break;
      case PJ_FEASIBLE: acc = infeasible(R, a); break;
      case PJ_FACE: {
        const bool f = face_of(R, a);
        gmem(g.face)[r] = f ? 1 : 0;
        acc = f ? 1.0 : 0.0;
        break;
      }
      case PJ_PHI: st3(g.out, r, phi_of(R, gmem(g.face)[r] != 0, a)); break;
      case PJ_BETA: st3(g.out, r, beta_of(R, gmem(g.face)[r] != 0, a, ld3(g.b, r))); break;
      case PJ_PHITPHI: {
        const V3 ph = ld3(g.b, r);
        const V3 px = proj(R, axmy(a, g.s0, ph));
        acc = dot3(sub(a, px), ph);
        break;
      }
      case PJ_STEP_LIMIT:
        if (R.has) {
          const double jx = dot3(R.n, a), jd = dot3(R.n, ld3(g.b, r));
          if (jd > kEps && jx > R.c) acc = (jx - R.c) / jd;
        }
        break;
      case PJ_PRECOND: {
        Vecs<double> v{};
        v.dinv = g.dinv;
        st3(g.out, r, precond_of(v, g.s0 != 0.0, R, gmem(g.face)[r] != 0, r, a, ld3(g.b, r)));
        break;
      }
      default: break;
    }
  }
  if (!g.part) return;  // uniform: no op that reduces passes a null part
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] = is_min ? fmin(red[threadIdx.x], red[threadIdx.x + w]) : red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) gmem(g.part)[blockIdx.x] = red[0];
}

// part[0..n) to out[0] in a fixed order (one workgroup).
__global__ void __launch_bounds__(256) reduce_parts(const double *__restrict__ part, int n, int is_min,
                                                    double *__restrict__ out) {
  __shared__ double red[256];
  double acc = is_min ? kDblMax : 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc = is_min ? fmin(acc, part[i]) : acc + part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      red[threadIdx.x] = is_min ? fmin(red[threadIdx.x], red[threadIdx.x + w]) : red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// reduce_parts (sums) of k partial arrays of n, one workgroup each.
__global__ void __launch_bounds__(256) reduce_parts_n(const double *__restrict__ part, int n,
                                                      double *__restrict__ out) {
  __shared__ double red[256];
  const double *__restrict__ p = part + (size_t)blockIdx.x * n;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc = acc + p[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Coupled-row solve (mprgp_general.cpp): CSR products one thread per row,
// terms in stored order; vector updates.
__global__ void __launch_bounds__(256) csr_mul(int rows, const int *__restrict__ ptr, const int *__restrict__ col,
                                               const double *__restrict__ val, const double *__restrict__ x,
                                               double *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
  out[i] = s;
}
// Several products with one CSR matrix: vector blockIdx.y, each row formed
// as csr_mul forms it.
struct CsrVecs {
  const double *x[kMaxCsrVecs];
  double *out[kMaxCsrVecs];
};
__global__ void __launch_bounds__(256) csr_mul_n(int rows, const int *__restrict__ ptr, const int *__restrict__ col,
                                                 const double *__restrict__ val, CsrVecs v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  const double *__restrict__ x = v.x[blockIdx.y];
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
  v.out[blockIdx.y][i] = s;
}
// out = a x + b y
__global__ void __launch_bounds__(256) axpby(size_t n, double a, const double *__restrict__ x, double b,
                                             const double *__restrict__ y, double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}
// out = x y (elementwise)
__global__ void __launch_bounds__(256) vmul(size_t n, const double *__restrict__ x, const double *__restrict__ y,
                                            double *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = x[i] * y[i];
}

// The power method's start vector (power_start_raw), divided by its norm.
__global__ void __launch_bounds__(256) power_start(double *__restrict__ out, size_t n, double nr) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double e = power_start_raw(i);
  out[i] = nr > 0.0 ? e / nr : e;
}

// FP32 copies for the mixed form.
__global__ void to_u16(const int *__restrict__ in, unsigned short *__restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (unsigned short)in[i];
}

__global__ void to_f32(const double *__restrict__ in, float *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (float)in[i];
}

// Streaming copy, one 16 B element per lane: the bandwidth probe that
// calibrates the roofline and the profiler's HBM byte counters.  (One element
// per lane over many workgroups measured faster on MI355X, 6.2 TB/s, than four
// elements per lane, 5.6 TB/s.)
__global__ void __launch_bounds__(256) copy_stream(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                   size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// The two state copies (1024 bytes) of QP i between a staging buffer and the
// QP's own state: one workgroup of 64 lanes per QP, one 16-byte word a lane.
static_assert(2 * sizeof(QpState) == 64 * sizeof(double2), "two states: 64 words of 16 bytes");
__global__ void __launch_bounds__(64) scatter_states(const double2 *__restrict__ src,
                                                     double2 *const *__restrict__ dst, int n) {
  if ((int)blockIdx.x < n) dst[blockIdx.x][threadIdx.x] = src[(size_t)blockIdx.x * 64 + threadIdx.x];
}
__global__ void __launch_bounds__(64) gather_states(const double2 *const *__restrict__ src,
                                                    double2 *__restrict__ dst, int n) {
  if ((int)blockIdx.x < n) dst[(size_t)blockIdx.x * 64 + threadIdx.x] = src[blockIdx.x][threadIdx.x];
}

// Plain SELL SpMV y = A x (used by tests and by the assembly's A x1 term).
__global__ void __launch_bounds__(256) spmv_plain(const int *__restrict__ sl_off,
                                                  const int *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const double *__restrict__ x, double *__restrict__ y,
                                                  int n_nodes, int accumulate) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  int s = r / kSlice, lane = r % kSlice;
  if (s * kSlice >= n_nodes) return;
  V3 acc[1];
  spmv_row<1>(sl_off, col, val, s, lane, [&](int j, V3 (&xs)[1]) { xs[0] = ld3(x, j); }, acc);
  if (r < n_nodes) {
    if (accumulate) acc[0] = add(acc[0], ld3(y, r));
    st3(y, r, acc[0]);
  }
}

// Host-callable launch wrappers ------------------------------------------

template <bool MX, int UB, bool NT, int SPW, bool C16>
static hipError_t launch_step_t(int ts, int grid, hipStream_t s, const QpDev *qps, const Tile *tiles,
                                int *running, int parity) {
  switch (ts) {
    case 1: mprgp_step<1, MX, UB, NT, SPW, C16><<<grid, 1 * kSlice, 0, s>>>(qps, tiles, running, parity); break;
    case 2: mprgp_step<2, MX, UB, NT, SPW, C16><<<grid, 2 * kSlice, 0, s>>>(qps, tiles, running, parity); break;
    case 4: mprgp_step<4, MX, UB, NT, SPW, C16><<<grid, 4 * kSlice, 0, s>>>(qps, tiles, running, parity); break;
    case 8: mprgp_step<8, MX, UB, NT, SPW, C16><<<grid, 8 * kSlice, 0, s>>>(qps, tiles, running, parity); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int SPW, bool C16>
static hipError_t launch_step_v(int ts, int grid, hipStream_t s, const QpDev *qps, const Tile *tiles,
                                int *running, int parity, bool mixed, int variant) {
  switch (variant) {
    case kStepCached:
      return mixed ? launch_step_t<true, kUnroll, false, SPW, C16>(ts, grid, s, qps, tiles, running, parity)
                   : launch_step_t<false, kUnroll, false, SPW, C16>(ts, grid, s, qps, tiles, running, parity);
    case kStepStream:
      return mixed ? launch_step_t<true, kUnroll, true, SPW, C16>(ts, grid, s, qps, tiles, running, parity)
                   : launch_step_t<false, kUnroll, true, SPW, C16>(ts, grid, s, qps, tiles, running, parity);
    case kStepTail:
      if constexpr (C16) {
        return hipErrorInvalidValue;  // the tail form reads 32-bit columns (launch_step)
      } else {
        return mixed ? launch_step_t<true, kUnrollTail, false, SPW, C16>(ts, grid, s, qps, tiles, running, parity)
                     : launch_step_t<false, kUnrollTail, false, SPW, C16>(ts, grid, s, qps, tiles, running, parity);
      }
    default: return hipErrorInvalidValue;
  }
}

// The 16-bit columns pay where a launch streams its matrices (2 of the 40
// bytes of an FP32 block); the tail form, bound by a lone row's chain of
// load rounds, keeps the 32-bit ones (the plain one-QP loop ran 2.4% slower
// with 16-bit ones).
hipError_t launch_step(int ts, int grid, hipStream_t s, const QpDev *qps, const Tile *tiles,
                       int *running, int parity, bool mixed, int variant, int spw, bool c16) {
  if (spw == 2) return launch_step_v<2, false>(ts, grid, s, qps, tiles, running, parity, mixed, variant);
  if (spw != 1) return hipErrorInvalidValue;
  return c16 && variant != kStepTail
             ? launch_step_v<1, true>(ts, grid, s, qps, tiles, running, parity, mixed, variant)
             : launch_step_v<1, false>(ts, grid, s, qps, tiles, running, parity, mixed, variant);
}

hipError_t launch_setup_rows_batch(hipStream_t s, const QpDev *qps, int count, int max_nodes) {
  if (count <= 0 || max_nodes <= 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  setup_rows_batch<<<dim3((max_nodes + 255) / 256, count), 256, 0, s>>>(qps);
  return hipGetLastError();
}
hipError_t launch_scatter_states(hipStream_t s, const QpState *src, QpState *const *dst, int n) {
  if (n <= 0) return hipSuccess;
  scatter_states<<<n, 64, 0, s>>>(reinterpret_cast<const double2 *>(src),
                                  reinterpret_cast<double2 *const *>(dst), n);
  return hipGetLastError();
}
hipError_t launch_gather_states(hipStream_t s, QpState *const *src, QpState *dst, int n) {
  if (n <= 0) return hipSuccess;
  gather_states<<<n, 64, 0, s>>>(reinterpret_cast<const double2 *const *>(src), reinterpret_cast<double2 *>(dst), n);
  return hipGetLastError();
}
hipError_t launch_gather_nodes(hipStream_t s, const double *in, double *out, const int *perm, int n) {
  if (n <= 0) return hipSuccess;
  gather_nodes<<<(n + 255) / 256, 256, 0, s>>>(in, out, perm, n);
  return hipGetLastError();
}
hipError_t launch_scatter_nodes(hipStream_t s, const double *in, double *out, const int *perm, int n) {
  if (n <= 0) return hipSuccess;
  scatter_nodes<<<(n + 255) / 256, 256, 0, s>>>(in, out, perm, n);
  return hipGetLastError();
}
hipError_t launch_setup_diag(hipStream_t s, const int *sl_off, const int *col, const double *val,
                             double *dinv, int n) {
  if (n <= 0) return hipSuccess;
  setup_diag<<<(n + 255) / 256, 256, 0, s>>>(sl_off, col, val, dinv, n);
  return hipGetLastError();
}
hipError_t launch_setup_rows(hipStream_t s, const double *dinv, const double *cn,
                             const unsigned char *cf, double *cs, int n) {
  if (n <= 0) return hipSuccess;
  setup_rows<<<(n + 255) / 256, 256, 0, s>>>(dinv, cn, cf, cs, n);
  return hipGetLastError();
}
hipError_t launch_collide_planes(hipStream_t s, const double *x, const double *planes, int n_planes,
                                 double *cn, double *cc, double *cj, unsigned char *cf, int *count, int n) {
  if (n <= 0) return hipSuccess;
  collide_planes<<<(n + 255) / 256, 256, 0, s>>>(x, planes, n_planes, cn, cc, cj, cf, count, n);
  return hipGetLastError();
}
hipError_t launch_get_diag(hipStream_t s, const int *sl_off, const int *col, const double *val, double *diag,
                           int n) {
  if (n <= 0) return hipSuccess;
  get_diag<<<(n + 255) / 256, 256, 0, s>>>(sl_off, col, val, diag, n);
  return hipGetLastError();
}
hipError_t launch_dot(hipStream_t s, const double *a, const double *b, size_t n, double *part, double *out) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  dot_parts<<<(unsigned)blocks, 256, 0, s>>>(a, b, n, part);
  reduce_parts<<<1, 256, 0, s>>>(part, (int)blocks, 0, out);
  return hipGetLastError();
}
hipError_t launch_dots(hipStream_t s, int k, const double *const *a, const double *const *b, size_t n,
                       double *part, double *out) {
  if (k <= 0 || n == 0) return hipSuccess;
  if (k > kMaxDots) return hipErrorInvalidValue;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  DotPairs pr{};
  for (int j = 0; j < k; ++j) {
    pr.a[j] = a[j];
    pr.b[j] = b[j];
  }
  dot_parts_n<<<dim3((unsigned)blocks, (unsigned)k), 256, 0, s>>>(pr, n, part);
  reduce_parts_n<<<(unsigned)k, 256, 0, s>>>(part, (int)blocks, out);
  return hipGetLastError();
}
hipError_t launch_projector(hipStream_t s, const ProjArgs &a, double *result) {
  if (a.n_nodes <= 0) return hipSuccess;
  const int blocks = (a.n_nodes + 255) / 256;
  projector_op<<<blocks, 256, 0, s>>>(a);
  if (a.part) reduce_parts<<<1, 256, 0, s>>>(a.part, blocks, a.op == PJ_STEP_LIMIT ? 1 : 0, result);
  return hipGetLastError();
}
hipError_t launch_csr_mul(hipStream_t s, int rows, const int *ptr, const int *col, const double *val,
                          const double *x, double *out) {
  if (rows <= 0) return hipSuccess;
  csr_mul<<<(rows + 255) / 256, 256, 0, s>>>(rows, ptr, col, val, x, out);
  return hipGetLastError();
}
hipError_t launch_csr_mul_n(hipStream_t s, int rows, const int *ptr, const int *col, const double *val, int k,
                            const double *const *x, double *const *out) {
  if (rows <= 0 || k <= 0) return hipSuccess;
  if (k > kMaxCsrVecs) return hipErrorInvalidValue;
  CsrVecs v{};
  for (int j = 0; j < k; ++j) {
    v.x[j] = x[j];
    v.out[j] = out[j];
  }
  csr_mul_n<<<dim3((rows + 255) / 256, (unsigned)k), 256, 0, s>>>(rows, ptr, col, val, v);
  return hipGetLastError();
}
hipError_t launch_axpby(hipStream_t s, size_t n, double a, const double *x, double b, const double *y, double *out) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  axpby<<<(unsigned)blocks, 256, 0, s>>>(n, a, x, b, y, out);
  return hipGetLastError();
}
hipError_t launch_vmul(hipStream_t s, size_t n, const double *x, const double *y, double *out) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  vmul<<<(unsigned)blocks, 256, 0, s>>>(n, x, y, out);
  return hipGetLastError();
}
hipError_t launch_power_start(hipStream_t s, double *out, size_t n, double nr) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  power_start<<<(unsigned)blocks, 256, 0, s>>>(out, n, nr);
  return hipGetLastError();
}
hipError_t launch_to_f32(hipStream_t s, const double *in, float *out, size_t n) {
  if (n == 0) return hipSuccess;
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  to_f32<<<(unsigned)blocks, 256, 0, s>>>(in, out, n);
  return hipGetLastError();
}
hipError_t launch_to_u16(hipStream_t s, const int *in, unsigned short *out, size_t n) {
  if (n == 0) return hipSuccess;
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  to_u16<<<(unsigned)blocks, 256, 0, s>>>(in, out, n);
  return hipGetLastError();
}
hipError_t launch_copy_stream(hipStream_t s, const void *in, void *out, size_t bytes) {
  const size_t n = bytes / sizeof(double2);
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  copy_stream<<<(unsigned)blocks, 256, 0, s>>>((const double2 *)in, (double2 *)out, n);
  return hipGetLastError();
}
hipError_t launch_spmv_plain(hipStream_t s, const int *sl_off, const int *col, const double *val,
                             const double *x, double *y, int n, int accumulate) {
  if (n <= 0) return hipSuccess;
  int slices = (n + kSlice - 1) / kSlice;
  int threads = slices * kSlice;
  spmv_plain<<<(threads + 255) / 256, 256, 0, s>>>(sl_off, col, val, x, y, n, accumulate);
  return hipGetLastError();
}

}  // namespace mprgp
