// mprgp_fem.hip -- device assembly of the contact QP's system,
// A = M + h^2 K(x_n) and b (FEMSystem::buildSystem, FEMSystem.cpp:167-192, and
// MprgpFemSolver::buildLinearSystem, MprgpFemSolver.cpp:315-356), for LINEAR
// and COROTATIONAL tets (FEMEnergy.cpp:657-1057).
//
// Three passes, no atomics, every sum in tet order so the result is
// deterministic and independent of the launch shape:
//   1. fem_element: one thread per tet computes h^2 K_e (16 blocks of 3x3)
//      and h^2 f_e (4 x 3) into a scratch array laid out [entry][tet] so the
//      per-thread writes are coalesced.
//   2. fem_blocks: one thread per block of A sums the contributions of the
//      tets sharing that node pair (a precomputed list in tet order), adds
//      the lumped mass on diagonal blocks and stores the block straight into
//      its SELL-64 slot.
//   3. fem_rhs: one thread per node sums its tets' forces, gravity and the
//      inertia term; A x_n is then added by the SELL SpMV.
//
// Element algebra.  With d = Dm^-1 (buildD, FEMMesh.cpp:166-184) the shape
// gradients are B_I = d^T e_I (I = 1..3) and B_0 = -(B_1+B_2+B_3), so
// F = sum_I x_I B_I^T (buildF, FEMMesh.cpp:123-157) and the 9x12 map of
// calcGComp3D (FEMUtils.h:30-41) acts blockwise.  For
// E = V (lambda/2 tr(e)^2 + mu e:e), e = sym(R^T F) - I (R = I for LINEAR;
// for COROT R is the polar rotation of F, held fixed in the Hessian):
//   f_I  = -V R (lambda tr(e) I + 2 mu e) B_I
//   K_IJ =  V [ lambda u_I u_J^T + mu ((R R^T)(B_I.B_J) + u_J u_I^T) ],  u_I = R B_I
// The polar rotation follows adjustF3D (FEMRotationUtil.h:959-1010): a cyclic
// Jacobi eigen-decomposition of F^T F, V with det +1, U = F V S^-1 with the
// column of the smallest singular value flipped when det U < 0, R = U V^T.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mprgp_host.h"

#pragma clang fp contract(off)

namespace mprgp {

// per tet, contiguous (tet-major, so a block's 9 values are one read for the
// kernels that sum them): kc K_e blocks, then h^2 f_e, then K_e c_e (the
// damping term)
constexpr int kElemOut = 16 * 9 + 12 + 12;

__device__ inline double det3(const double *M) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
         M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// cyclic Jacobi on a symmetric 3x3 (row-major, destroyed); ascending w, V columns
__device__ void jacobi3(double *A, double *w, double *V) {
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
        for (int k = 0; k < 3; ++k) {
          double akp = A[k * 3 + p], akq = A[k * 3 + q];
          A[k * 3 + p] = c * akp - s * akq;
          A[k * 3 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = A[p * 3 + k], aqk = A[q * 3 + k];
          A[p * 3 + k] = c * apk - s * aqk;
          A[q * 3 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = V[k * 3 + p], vkq = V[k * 3 + q];
          V[k * 3 + p] = c * vkp - s * vkq;
          V[k * 3 + q] = s * vkp + c * vkq;
        }
      }
  }
  w[0] = A[0];
  w[1] = A[4];
  w[2] = A[8];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2 - i; ++j)
      if (w[j] > w[j + 1]) {
        double t = w[j];
        w[j] = w[j + 1];
        w[j + 1] = t;
        for (int k = 0; k < 3; ++k) {
          double tv = V[k * 3 + j];
          V[k * 3 + j] = V[k * 3 + j + 1];
          V[k * 3 + j + 1] = tv;
        }
      }
}

__device__ void polar_rotation(const double *F, double *R) {
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
    for (int k = 0; k < 3; ++k) V[k * 3] = -V[k * 3];
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
      double a[3] = {U[good[0]], U[3 + good[0]], U[6 + good[0]]};
      double b[3];
      if (nr == 1) {  // degenerate tet: cross with the axis least aligned with a
        int ax = 0;
        if (fabs(a[1]) < fabs(a[ax])) ax = 1;
        if (fabs(a[2]) < fabs(a[ax])) ax = 2;
        b[0] = b[1] = b[2] = 0.0;
        b[ax] = 1.0;
      } else {
        b[0] = U[good[1]];
        b[1] = U[3 + good[1]];
        b[2] = U[6 + good[1]];
      }
      double cx[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      if (nr == 1) {
        double nd = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]);
        cx[0] /= nd;
        cx[1] /= nd;
        cx[2] /= nd;
      }
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

// rot: per-tet rotations, [entry][tet].  corot && !lag: R is the polar
// rotation of F and is stored; corot && lag: R is the stored rotation of the
// last full assembly (the element is then linear in x with exactly the
// stiffness of that assembly).  rhs_only skips the stiffness blocks.
// Outputs (FEMSystem::buildSystem, FEMSystem.cpp:167-192): the stiffness
// blocks times kc = h^2 + h alpha_m, the force times h^2, and with damp set
// the unscaled K_e c_e for c = (x_n - x1) h (the stiffness-damping term, CRHS).
__global__ void __launch_bounds__(128) fem_element(int T, const int4 *__restrict__ tets,
                                                   const double *__restrict__ dinv,
                                                   const double *__restrict__ vol,
                                                   const double *__restrict__ lam_mu,
                                                   const double *__restrict__ x,
                                                   const double *__restrict__ xn, int corot, int lag,
                                                   int rhs_only, int damp, double kc, double h2, double h,
                                                   double *__restrict__ rot, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int4 v = tets[t];
  const int vi[4] = {v.x, v.y, v.z, v.w};
  double xe[12];
  for (int i = 0; i < 4; ++i)
    for (int r = 0; r < 3; ++r) xe[i * 3 + r] = x[3 * (size_t)vi[i] + r];
  double d[9];
  for (int k = 0; k < 9; ++k) d[k] = dinv[(size_t)k * T + t];
  // shape gradients B_I (I = 0..3): B_I[c] = d[I-1][c], B_0 = -sum
  double B[4][3];
  for (int c = 0; c < 3; ++c) {
    B[1][c] = d[0 * 3 + c];
    B[2][c] = d[1 * 3 + c];
    B[3][c] = d[2 * 3 + c];
    B[0][c] = -((d[0 * 3 + c] + d[1 * 3 + c]) + d[2 * 3 + c]);
  }
  double F[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += (xe[(i + 1) * 3 + r] - xe[r]) * d[i * 3 + c];
      F[r * 3 + c] = s;
    }
  double R[9];
  if (corot && lag) {
    for (int k = 0; k < 9; ++k) R[k] = rot[(size_t)k * T + t];
  } else if (corot) {
    polar_rotation(F, R);
    for (int k = 0; k < 9; ++k) rot[(size_t)k * T + t] = R[k];
  } else {
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  }
  double S[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (corot) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += R[k * 3 + i] * F[k * 3 + j];
        S[i * 3 + j] = s;
      } else {
        S[i * 3 + j] = F[i * 3 + j];
      }
    }
  double eps[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) eps[i * 3 + j] = 0.5 * (S[i * 3 + j] + S[j * 3 + i]) - (i == j ? 1.0 : 0.0);
  const double lam = lam_mu[t], mu = lam_mu[(size_t)T + t], V = vol[t];
  const double tr = eps[0] + eps[4] + eps[8];
  double P[9], RP[9], RRt[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P[i * 3 + j] = lam * tr * (i == j ? 1.0 : 0.0) + 2.0 * mu * eps[i * 3 + j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0, s2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        s += R[i * 3 + k] * P[k * 3 + j];
        s2 += R[i * 3 + k] * R[j * 3 + k];
      }
      RP[i * 3 + j] = s;
      RRt[i * 3 + j] = s2;
    }
  double u[4][3];
  for (int I = 0; I < 4; ++I)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += R[r * 3 + c] * B[I][c];
      u[I][r] = s;
    }
  // kc K_e blocks; K_e c_e for the damping term
  double ce[12];
  if (damp)
    for (int i = 0; i < 4; ++i)
      for (int r = 0; r < 3; ++r) ce[i * 3 + r] = (xn[3 * (size_t)vi[i] + r] - xe[i * 3 + r]) * h;
  for (int I = 0; I < 4 && (!rhs_only || damp); ++I) {
    double kd[3] = {0.0, 0.0, 0.0};
    for (int J = 0; J < 4; ++J) {
      const double bb = (B[I][0] * B[J][0] + B[I][1] * B[J][1]) + B[I][2] * B[J][2];
      for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
          double k = V * (lam * u[I][r] * u[J][s] + mu * (RRt[r * 3 + s] * bb + u[J][r] * u[I][s]));
          if (!rhs_only) out[(size_t)t * kElemOut + ((I * 4 + J) * 9) + r * 3 + s] = k * kc;
          if (damp) kd[r] += k * ce[J * 3 + s];
        }
    }
    if (damp)
      for (int r = 0; r < 3; ++r) out[(size_t)t * kElemOut + 156 + I * 3 + r] = kd[r];
  }
  // h^2 f_e
  for (int I = 0; I < 4; ++I)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += RP[r * 3 + c] * B[I][c];
      out[(size_t)t * kElemOut + 144 + I * 3 + r] = -(V * s) * h2;
    }
}

// One thread per block of A (node-order block index b): sums the tets'
// h^2 K_e blocks in tet order, adds M on the diagonal, stores the block in
// its SELL slot (value base vbase[b], entry planes 64 apart).
__global__ void fem_blocks(long long nb, int T, const int *__restrict__ ptr, const int *__restrict__ ctr,
                           const int *__restrict__ vbase, const int *__restrict__ diag_node,
                           const double *__restrict__ mass, double mc, const double *__restrict__ elem,
                           double *__restrict__ val) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double acc[9];
  for (int e = 0; e < 9; ++e) acc[e] = 0.0;
  for (int k = ptr[b]; k < ptr[b + 1]; ++k) {
    const int code = ctr[k];
    const int t = code >> 4, IJ = code & 15;
    const double *eb = elem + (size_t)t * kElemOut + IJ * 9;  // the block's 9 values, contiguous
    for (int e = 0; e < 9; ++e) acc[e] += eb[e];
  }
  const int dn = diag_node[b];
  if (dn >= 0) {
    const double M = mass[dn] * mc;  // M (1 + h alpha_k)
    acc[0] += M;
    acc[4] += M;
    acc[8] += M;
  }
  double *o = val + vbase[b];
  for (int e = 0; e < 9; ++e) o[(size_t)e * kSlice] = acc[e];
}

// b (node order, before A x1): sum of h^2 f_e(x1) in tet order, then the
// stiffness damping alpha_m K c, then gravity h^2 rho m g, then
// M ((x_n + h v_n - x1) + c alpha_k) with c = (x_n - x1) h  (FEMSystem.cpp:176-186
// with MRHSL = PSI - X1 and CRHS = c, MprgpFemSolver.cpp:336-340).  x1 is the
// Newton iterate (x_n on the first iteration, where c = 0).
__global__ void fem_rhs(int N, int T, const int *__restrict__ ptr, const int *__restrict__ ctr,
                        const double *__restrict__ node_mass, double density,
                        const double *__restrict__ elem, const double *__restrict__ x,
                        const double *__restrict__ v, const double *__restrict__ x1, double g0, double g1,
                        double g2, double h, int damp, double alpha_k, double alpha_m,
                        double *__restrict__ b) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
    const int code = ctr[k];
    const int t = code >> 2, I = code & 3;
    for (int r = 0; r < 3; ++r) s[r] += elem[(size_t)t * kElemOut + 144 + I * 3 + r];
  }
  if (damp) {
    double kd[3] = {0.0, 0.0, 0.0};
    for (int k = ptr[a]; k < ptr[a + 1]; ++k) {
      const int code = ctr[k];
      const int t = code >> 2, I = code & 3;
      for (int r = 0; r < 3; ++r) kd[r] += elem[(size_t)t * kElemOut + 156 + I * 3 + r];
    }
    for (int r = 0; r < 3; ++r) s[r] += kd[r] * alpha_m;
  }
  const double g[3] = {g0, g1, g2};
  const double m = node_mass[a];
  const double M = m * density;
  const double h2 = h * h;
  for (int r = 0; r < 3; ++r) s[r] += ((g[r] * density) * m) * h2;
  for (int r = 0; r < 3; ++r) {
    const double xn = x[3 * (size_t)a + r], y = x1[3 * (size_t)a + r];
    const double psi = xn + h * v[3 * (size_t)a + r];
    s[r] += M * ((psi - y) + ((xn - y) * h) * alpha_k);
  }
  for (int r = 0; r < 3; ++r) b[3 * (size_t)a + r] = s[r];
}

// max_i |a_i - b_i| (updatePos's tol_err, MprgpFemSolver.cpp:368-374) in
// two kernels: each workgroup's maximum over a strided share, then the
// maximum of those.  fmax is exact and ignores NaN in any order, so the
// result does not depend on the split.
constexpr int kMaxParts = 256;
__global__ void __launch_bounds__(256) max_abs_diff_parts(const double *__restrict__ a,
                                                          const double *__restrict__ b, size_t n,
                                                          double *__restrict__ part) {
  __shared__ double red[256];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    m = fmax(m, fabs(a[i] - b[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(256) max_parts(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double red[256];
  red[threadIdx.x] = (int)threadIdx.x < np ? part[threadIdx.x] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}
// out[0] = max |a - b|, out[1 .. kMaxParts] the partial maxima
hipError_t launch_max_abs_diff(hipStream_t s, const double *a, const double *b, size_t n, double *out) {
  const int np = (int)std::min<size_t>(kMaxParts, std::max<size_t>(1, (n + 4095) / 4096));
  max_abs_diff_parts<<<np, 256, 0, s>>>(a, b, n, out + 1);
  max_parts<<<1, 256, 0, s>>>(out + 1, np, out);
  return hipGetLastError();
}

// updateMesh (MprgpFemSolver.cpp:377-390) with initVel's PHI and PSI
// (:295-313) under implicit Euler (beta = 1, beta2 = 0.5, gamma = 1,
// FEMSystem.h:140-142): the previous acceleration enters with the factors
// 1 - gamma = 0 and 1 - 2 beta2 = 0, so a = (x - (x_n + h v_n)) / (h h) and
// v = v_n + a h.  Then x_n = x, v_n = v.
__global__ void euler_update(size_t n, const double *__restrict__ x, double *__restrict__ xn,
                             double *__restrict__ vn, double h) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double psi = xn[i] + h * vn[i];
  const double acc = (x[i] - psi) / ((1.0 * h) * h);
  vn[i] = vn[i] + (acc * 1.0) * h;
  xn[i] = x[i];
}

}  // namespace mprgp

using namespace mprgp;

struct mprgp_fem {
  mprgp_ctx *ctx = nullptr;
  mprgp_matrix *A = nullptr;
  int N = 0, T = 0;
  long long nb = 0;
  void *dev = nullptr;  // one allocation for all static arrays
  size_t dev_bytes = 0;
  int4 *tets = nullptr;
  double *dinv = nullptr, *vol = nullptr, *lam_mu = nullptr, *node_mass = nullptr, *mass_rho = nullptr;
  int *blk_ptr = nullptr, *blk_ctr = nullptr, *blk_vbase = nullptr, *blk_diag = nullptr;
  int *node_ptr = nullptr, *node_ctr = nullptr;
  double *rot = nullptr;   // per-tet rotations of the last corotational assembly
  bool rot_valid = false;
  double density = 0.0;
  double min_mass = 0.0;  // smallest lumped node mass rho sum(V/4)
  double alpha_k = 0.0, alpha_m = 0.0;  // Rayleigh damping (FEMSystem::setDamping)
  // device-resident body state (node order): positions x_n, velocities v_n,
  // the Newton iterate x1 and the next one x2 (allocated by set_state)
  double *state = nullptr;
  double *xn = nullptr, *vn = nullptr, *x1 = nullptr, *x2 = nullptr;
  bool has_state = false;
  bool step_open = false;  // begin_step ran and end_step has not
};

namespace {

template <class T>
T *carve(char *&p, size_t count) {
  uintptr_t u = ((uintptr_t)p + 255) & ~(uintptr_t)255;
  T *r = (T *)u;
  p = (char *)(u + sizeof(T) * count);
  return r;
}

}  // namespace

extern "C" int mprgp_fem_create(mprgp_ctx *ctx, int N, int T, const int *tets_in, const double *X0,
                                const double *young, const double *poisson, double density, mprgp_fem **out) {
  MPRGP_TRY
  if (!ctx || !out || N <= 0 || T <= 0 || !tets_in || !X0 || !young || !poisson)
    return fail(MPRGP_EBADINPUT, "bad FEM arguments");
  if (!(density > 0.0) || !std::isfinite(density)) return fail(MPRGP_EBADINPUT, "density must be positive");
  if ((long long)T * 16 >= (1LL << 31)) return fail(MPRGP_EBADINPUT, "too many tets");
  *out = nullptr;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  std::vector<int> tets(tets_in, tets_in + 4 * (size_t)T);
  std::vector<double> dinv(9 * (size_t)T), vol(T), lam_mu(2 * (size_t)T), mass(N, 0.0);
  for (int t = 0; t < T; ++t) {
    int *v = &tets[4 * (size_t)t];
    for (int i = 0; i < 4; ++i)
      if (v[i] < 0 || v[i] >= N) return fail(MPRGP_EBADINPUT, "tet %d: node %d out of range", t, v[i]);
    auto vol6 = [&](const int *w) {
      double u[3], a[3], c[3];
      for (int r = 0; r < 3; ++r) {
        u[r] = X0[3 * (size_t)w[1] + r] - X0[3 * (size_t)w[0] + r];
        a[r] = X0[3 * (size_t)w[2] + r] - X0[3 * (size_t)w[0] + r];
        c[r] = X0[3 * (size_t)w[3] + r] - X0[3 * (size_t)w[0] + r];
      }
      double x0 = u[1] * a[2] - u[2] * a[1], x1 = u[2] * a[0] - u[0] * a[2], x2 = u[0] * a[1] - u[1] * a[0];
      return (x0 * c[0] + x1 * c[1] + x2 * c[2]) / 6.0;
    };
    if (vol6(v) < 0.0) std::swap(v[2], v[3]);  // makePositive (FEMMesh.cpp:158-165)
    vol[t] = vol6(v);
    if (!(vol[t] > 0.0)) return fail(MPRGP_EBADINPUT, "tet %d is degenerate", t);
    double Dm[9];
    for (int r = 0; r < 3; ++r)
      for (int i = 0; i < 3; ++i) Dm[r * 3 + i] = X0[3 * (size_t)v[i + 1] + r] - X0[3 * (size_t)v[0] + r];
    double c00 = Dm[4] * Dm[8] - Dm[5] * Dm[7], c01 = Dm[5] * Dm[6] - Dm[3] * Dm[8],
           c02 = Dm[3] * Dm[7] - Dm[4] * Dm[6];
    double id = 1.0 / (Dm[0] * c00 + Dm[1] * c01 + Dm[2] * c02);
    double I9[9] = {c00 * id, (Dm[2] * Dm[7] - Dm[1] * Dm[8]) * id, (Dm[1] * Dm[5] - Dm[2] * Dm[4]) * id,
                    c01 * id, (Dm[0] * Dm[8] - Dm[2] * Dm[6]) * id, (Dm[2] * Dm[3] - Dm[0] * Dm[5]) * id,
                    c02 * id, (Dm[1] * Dm[6] - Dm[0] * Dm[7]) * id, (Dm[0] * Dm[4] - Dm[1] * Dm[3]) * id};
    for (int k = 0; k < 9; ++k) dinv[(size_t)k * T + t] = I9[k];
    const double E = young[t], nu = poisson[t];
    if (!(E > 0.0) || !(nu > -1.0 && nu < 0.5)) return fail(MPRGP_EBADINPUT, "tet %d: bad material", t);
    lam_mu[(size_t)T + t] = E / (2.0 * (1.0 + nu));                 // setParam (FEMEnergy.cpp:383)
    lam_mu[t] = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));             // (FEMEnergy.cpp:384)
  }
  for (int t = 0; t < T; ++t)
    for (int j = 0; j < 4; ++j) mass[tets[4 * (size_t)t + j]] += vol[t] / 4.0;  // FEMMesh.cpp:507-509
  // node adjacency -> block pattern (sorted neighbours, self included)
  std::vector<std::vector<int>> nbr(N);
  for (int t = 0; t < T; ++t)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) nbr[tets[4 * (size_t)t + i]].push_back(tets[4 * (size_t)t + j]);
  std::vector<int> rowptr(N + 1, 0), colidx;
  for (int a = 0; a < N; ++a) {
    auto &v = nbr[a];
    v.push_back(a);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    rowptr[a + 1] = rowptr[a] + (int)v.size();
    colidx.insert(colidx.end(), v.begin(), v.end());
  }
  const long long nb = rowptr[N];
  mprgp_matrix *A = nullptr;
  if (int rc = build_matrix(ctx, N, rowptr.data(), colidx.data(), nullptr, &A)) return rc;
  // per-block contribution lists in tet order
  auto blk_of = [&](int a, int b) {
    auto it = std::lower_bound(colidx.begin() + rowptr[a], colidx.begin() + rowptr[a + 1], b);
    return (int)(it - colidx.begin());
  };
  std::vector<int> bcnt(nb + 1, 0), ncnt(N + 1, 0);
  for (int t = 0; t < T; ++t)
    for (int i = 0; i < 4; ++i) {
      ncnt[tets[4 * (size_t)t + i] + 1]++;
      for (int j = 0; j < 4; ++j) bcnt[blk_of(tets[4 * (size_t)t + i], tets[4 * (size_t)t + j]) + 1]++;
    }
  for (long long b = 0; b < nb; ++b) bcnt[b + 1] += bcnt[b];
  for (int a = 0; a < N; ++a) ncnt[a + 1] += ncnt[a];
  std::vector<int> bctr(bcnt[nb]), nctr(ncnt[N]);
  {
    std::vector<int> bf(bcnt.begin(), bcnt.end() - 1), nf(ncnt.begin(), ncnt.end() - 1);
    for (int t = 0; t < T; ++t)
      for (int i = 0; i < 4; ++i) {
        int a = tets[4 * (size_t)t + i];
        nctr[nf[a]++] = t * 4 + i;
        for (int j = 0; j < 4; ++j) bctr[bf[blk_of(a, tets[4 * (size_t)t + j])]++] = t * 16 + i * 4 + j;
      }
  }
  std::vector<int> vbase(nb), diag(nb, -1);
  for (int a = 0; a < N; ++a)
    for (int k = rowptr[a]; k < rowptr[a + 1]; ++k) {
      long long slot = A->blk_slot[k];
      int s = (int)(std::upper_bound(A->sl_off_h.begin(), A->sl_off_h.end(), (int)slot) - A->sl_off_h.begin()) - 1;
      long long b0 = A->sl_off_h[s];
      long long l = (slot - b0) % kSlice, kk = (slot - b0) / kSlice;
      vbase[k] = (int)(b0 * 9 + kk * 9 * kSlice + l);
      if (colidx[k] == a) diag[k] = a;
    }
  std::vector<double> mass_rho(N);
  for (int a = 0; a < N; ++a) mass_rho[a] = mass[a] * density;  // FEMSystem::onDirty (FEMSystem.cpp:561-562)
  mprgp_fem *f = new mprgp_fem();
  f->ctx = ctx;
  f->A = A;
  f->N = N;
  f->T = T;
  f->nb = nb;
  f->density = density;
  f->min_mass = *std::min_element(mass_rho.begin(), mass_rho.end());
  size_t bytes = 256 * 17 + sizeof(int4) * T + sizeof(double) * (9 + 1 + 2 + 9) * (size_t)T +
                 sizeof(double) * 2 * N + sizeof(int) * ((nb + 1) + bctr.size() + 2 * nb + (N + 1) + nctr.size());
  if (int rc = ctx->alloc(&f->dev, bytes)) {
    mprgp_matrix_destroy(A);
    delete f;
    return rc;
  }
  f->dev_bytes = bytes;
  char *p = (char *)f->dev;
  f->tets = carve<int4>(p, T);
  f->dinv = carve<double>(p, 9 * (size_t)T);
  f->vol = carve<double>(p, T);
  f->lam_mu = carve<double>(p, 2 * (size_t)T);
  f->node_mass = carve<double>(p, N);
  f->mass_rho = carve<double>(p, N);
  f->blk_ptr = carve<int>(p, nb + 1);
  f->blk_ctr = carve<int>(p, bctr.size());
  f->blk_vbase = carve<int>(p, nb);
  f->blk_diag = carve<int>(p, nb);
  f->node_ptr = carve<int>(p, N + 1);
  f->node_ctr = carve<int>(p, nctr.size());
  f->rot = carve<double>(p, 9 * (size_t)T);
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(f->tets, tets.data(), sizeof(int) * 4 * (size_t)T, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->dinv, dinv.data(), sizeof(double) * dinv.size(), hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->vol, vol.data(), sizeof(double) * T, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->lam_mu, lam_mu.data(), sizeof(double) * 2 * (size_t)T, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->node_mass, mass.data(), sizeof(double) * N, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->mass_rho, mass_rho.data(), sizeof(double) * N, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->blk_ptr, bcnt.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->blk_ctr, bctr.data(), sizeof(int) * bctr.size(), hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->blk_vbase, vbase.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->blk_diag, diag.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->node_ptr, ncnt.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, s);
  if (!e) e = hipMemcpyAsync(f->node_ctr, nctr.data(), sizeof(int) * nctr.size(), hipMemcpyHostToDevice, s);
  if (!e) e = hipStreamSynchronize(s);
  if (e) {
    mprgp_fem_destroy(f);
    return fail(MPRGP_EHIP, "FEM upload failed: %s", hipGetErrorString(e));
  }
  *out = f;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" mprgp_matrix *mprgp_fem_matrix(mprgp_fem *f) { return f ? f->A : nullptr; }

namespace {

int check_material(const mprgp_fem *f, int material) {
  const bool rhs_only = (material & MPRGP_RHS_ONLY) != 0;
  material &= ~MPRGP_RHS_ONLY;
  if (material != MPRGP_LINEAR && material != MPRGP_COROTATIONAL)
    return fail(MPRGP_EBADINPUT, "unknown material %d", material);
  if (rhs_only && f->A->version == 0) return fail(MPRGP_ESTATE, "RHS-only assembly before any full assembly");
  if (rhs_only && material == MPRGP_COROTATIONAL && !f->rot_valid)
    return fail(MPRGP_ESTATE, "corotational RHS-only assembly needs a full corotational assembly first");
  return MPRGP_OK;
}

int check_finite(const double *v, size_t n, const char *what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return fail(MPRGP_EBADINPUT, "non-finite %s", what);
  return MPRGP_OK;
}

// A (unless RHS_ONLY) and b at the Newton iterate x1 from device arrays in
// node order (buildLinearSystem, MprgpFemSolver.cpp:315-356).  The caller holds
// the context lock, has checked the material and keeps the arrays alive.
int assemble_dev(mprgp_fem *f, int material, const double *xn, const double *vn, const double *x1,
                 const double *gravity, double h, mprgp_qp *qp) {
  const bool rhs_only = (material & MPRGP_RHS_ONLY) != 0;
  material &= ~MPRGP_RHS_ONLY;
  mprgp_ctx *ctx = f->ctx;
  const size_t need = sizeof(double) * ((size_t)kElemOut * f->T + 3 * (size_t)f->N) + 1024;
  if (int rc = ctx->scratch_reserve(need)) return rc;
  const int N = f->N, T = f->T;
  double *elem = (double *)ctx->scratch;
  double *db = elem + (size_t)kElemOut * T;
  hipStream_t s = ctx->stream;
  // buildLinearSystem's coefficients with beta = 1 (MprgpFemSolver.cpp:331-339):
  // bdt = h, KCoef = FCoef = bdt h, CCoef = bdt, MCoef = 1
  const double bdt = 1.0 * h;
  const double h2 = bdt * h;
  const double kc = h2 + bdt * f->alpha_m;
  const double mc = 1.0 + bdt * f->alpha_k;
  const int damp = f->alpha_m != 0.0 ? 1 : 0;
  const int corot = material == MPRGP_COROTATIONAL;
  fem_element<<<(T + 127) / 128, 128, 0, s>>>(T, f->tets, f->dinv, f->vol, f->lam_mu, x1, xn, corot,
                                              rhs_only ? 1 : 0, rhs_only ? 1 : 0, damp, kc, h2, bdt, f->rot,
                                              elem);
  MPRGP_HIP(hipGetLastError());
  if (!rhs_only) {
    fem_blocks<<<(unsigned)((f->nb + 255) / 256), 256, 0, s>>>(f->nb, T, f->blk_ptr, f->blk_ctr, f->blk_vbase,
                                                                f->blk_diag, f->mass_rho, mc, elem, f->A->d_val);
    MPRGP_HIP(hipGetLastError());
  }
  fem_rhs<<<(N + 255) / 256, 256, 0, s>>>(N, T, f->node_ptr, f->node_ctr, f->node_mass, f->density, elem, xn,
                                          vn, x1, gravity[0], gravity[1], gravity[2], bdt, damp, f->alpha_k,
                                          f->alpha_m, db);
  MPRGP_HIP(hipGetLastError());
  // b += A x1 (MprgpFemSolver.cpp:349), in the QP's internal node order
  MPRGP_HIP(launch_gather_nodes(s, db, qp->b, f->A->d_perm, N));
  MPRGP_HIP(launch_gather_nodes(s, x1, qp->stage, f->A->d_perm, N));
  MPRGP_HIP(launch_spmv_plain(s, f->A->d_sl_off, f->A->d_col, f->A->d_val, qp->stage, qp->b, N, 1));
  if (!rhs_only) MPRGP_HIP(launch_setup_diag(s, f->A->d_sl_off, f->A->d_col, f->A->d_val, f->A->d_dinv, N));
  MPRGP_HIP(hipStreamSynchronize(s));
  if (!rhs_only) {
    f->A->version++;
    f->rot_valid = corot;
    // A = M (1 + h alpha_k) + K (h^2 + h alpha_m), K positive semidefinite
    // (LINEAR, or COROTATIONAL with the rotations held): its smallest
    // eigenvalue is at least the smallest diagonal of the mass term
    f->A->lambda_min_lower = kc >= 0.0 && mc > 0.0 ? f->min_mass * mc : 0.0;
  }
  qp->has_rhs = true;
  return MPRGP_OK;
}

}  // namespace

extern "C" int mprgp_fem_assemble_at(mprgp_fem *f, int material, const double *x_n, const double *v_n,
                                     const double *x1, const double *gravity, double h, mprgp_qp *qp) {
  MPRGP_TRY
  if (!f || !x_n || !v_n || !gravity || !qp) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (int rc = check_material(f, material)) return rc;
  if (qp->A != f->A) return fail(MPRGP_EBADINPUT, "qp was not created on this FEM's matrix");
  MPRGP_NOT_BUSY(qp);
  if (!(h > 0.0) || !std::isfinite(h)) return fail(MPRGP_EBADINPUT, "h must be positive");
  const size_t n = 3 * (size_t)f->N;
  if (int rc = check_finite(x_n, n, "x_n")) return rc;
  if (int rc = check_finite(v_n, n, "v_n")) return rc;
  if (x1)
    if (int rc = check_finite(x1, n, "x1")) return rc;
  if (int rc = check_finite(gravity, 3, "gravity")) return rc;
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  // scratch: assemble_dev's element array and b, then the staged host arrays
  const size_t elem_b = (size_t)kElemOut * f->T + n;
  if (int rc = ctx->scratch_reserve(sizeof(double) * (elem_b + 3 * n) + 1024)) return rc;
  double *stage = (double *)ctx->scratch + elem_b;
  hipStream_t s = ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(stage, x_n, sizeof(double) * n, hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(stage + n, v_n, sizeof(double) * n, hipMemcpyHostToDevice, s));
  if (x1) MPRGP_HIP(hipMemcpyAsync(stage + 2 * n, x1, sizeof(double) * n, hipMemcpyHostToDevice, s));
  return assemble_dev(f, material, stage, stage + n, x1 ? stage + 2 * n : stage, gravity, h, qp);
  MPRGP_CATCH
}

extern "C" int mprgp_fem_assemble(mprgp_fem *f, int material, const double *x_n, const double *v_n,
                                  const double *gravity, double h, mprgp_qp *qp) {
  MPRGP_TRY
  return mprgp_fem_assemble_at(f, material, x_n, v_n, nullptr, gravity, h, qp);
  MPRGP_CATCH
}

extern "C" int mprgp_fem_set_damping(mprgp_fem *f, double alpha_k, double alpha_m) {
  MPRGP_TRY
  if (!f) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!(alpha_k >= 0.0) || !(alpha_m >= 0.0) || !std::isfinite(alpha_k) || !std::isfinite(alpha_m))
    return fail(MPRGP_EBADINPUT, "damping must be finite and >= 0");  // Simulator.cpp:162-163
  f->alpha_k = alpha_k;
  f->alpha_m = alpha_m;
  return MPRGP_OK;
  MPRGP_CATCH
}

// ---------------------------------------------------------------- time step

extern "C" int mprgp_fem_set_state(mprgp_fem *f, const double *x, const double *v) {
  MPRGP_TRY
  if (!f || !x || !v) return fail(MPRGP_EBADINPUT, "NULL argument");
  const size_t n = 3 * (size_t)f->N;
  if (int rc = check_finite(x, n, "x")) return rc;
  if (int rc = check_finite(v, n, "v")) return rc;
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  if (!f->state) {
    if (int rc = ctx->alloc((void **)&f->state, sizeof(double) * 4 * n)) return rc;
    f->xn = f->state;
    f->vn = f->xn + n;
    f->x1 = f->vn + n;
    f->x2 = f->x1 + n;
  }
  hipStream_t s = ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(f->xn, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(f->vn, v, sizeof(double) * n, hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  f->has_state = true;
  f->step_open = false;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_fem_get_state(mprgp_fem *f, double *x, double *v) {
  MPRGP_TRY
  if (!f) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!f->has_state) return fail(MPRGP_ESTATE, "no state was set");
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  const size_t n = 3 * (size_t)f->N;
  hipStream_t s = ctx->stream;
  if (x) MPRGP_HIP(hipMemcpyAsync(x, f->xn, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (v) MPRGP_HIP(hipMemcpyAsync(v, f->vn, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  return MPRGP_OK;
  MPRGP_CATCH
}

namespace {

int check_step(mprgp_fem *f, mprgp_qp *qp, const mprgp_step_params *sp) {
  if (!f || !qp || !sp) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (qp->A != f->A) return fail(MPRGP_EBADINPUT, "qp was not created on this FEM's matrix");
  if (!f->has_state) return fail(MPRGP_ESTATE, "no state was set");
  if (!(sp->h > 0.0) || !std::isfinite(sp->h)) return fail(MPRGP_EBADINPUT, "h must be positive");
  if (int rc = check_finite(sp->gravity, 3, "gravity")) return rc;
  if (sp->n_planes < 0 || (sp->n_planes > 0 && !sp->planes)) return fail(MPRGP_EBADINPUT, "bad planes");
  return check_material(f, sp->material);
}

}  // namespace

extern "C" int mprgp_fem_begin_step(mprgp_fem *f, mprgp_qp *qp, const mprgp_step_params *sp, int *rows) {
  MPRGP_TRY
  if (int rc = check_step(f, qp, sp)) return rc;
  MPRGP_NOT_BUSY(qp);
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  hipStream_t s = ctx->stream;
  const int N = f->N;
  const size_t n = 3 * (size_t)N;
  // initPos/initVel: x1 = x_n; the QP's iterate (and reset point) is x_n
  MPRGP_HIP(hipMemcpyAsync(f->x1, f->xn, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  MPRGP_HIP(launch_gather_nodes(s, f->xn, qp->x, f->A->d_perm, N));
  MPRGP_HIP(hipMemcpyAsync(qp->x0, qp->x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  qp->has_x = true;
  // handleCollDetection: the planes at x_n (DCDhandle, MprgpFemSolver.cpp:255-271)
  int m = 0;
  if (int rc = mprgp_qp_collide_planes(qp, sp->n_planes, sp->planes, &m)) return rc;
  if (rows) *rows = m;
  // the first buildLinearSystem of forward's Newton loop, at x1 = x_n
  if (int rc = assemble_dev(f, sp->material, f->xn, f->vn, f->x1, sp->gravity, sp->h, qp)) return rc;
  f->step_open = true;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_fem_end_step(mprgp_fem *f, mprgp_qp *qp, const mprgp_step_params *sp, double *delta) {
  MPRGP_TRY
  if (int rc = check_step(f, qp, sp)) return rc;
  if (!f->step_open) return fail(MPRGP_ESTATE, "no step was begun");
  MPRGP_NOT_BUSY(qp);
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  hipStream_t s = ctx->stream;
  const int N = f->N;
  const size_t n = 3 * (size_t)N;
  // updatePos: x2 = the solve's x (node order), tol_err = max |x2 - x1|
  MPRGP_HIP(launch_scatter_nodes(s, qp->x, f->x2, f->A->d_perm, N));
  if (delta) {
    if (int rc = ctx->scratch_reserve(sizeof(double) * (kMaxParts + 1))) return rc;
    MPRGP_HIP(launch_max_abs_diff(s, f->x2, f->x1, n, (double *)ctx->scratch));
    MPRGP_HIP(hipMemcpyAsync(delta, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  // updateMesh: velocities from x2, then x_n = x2
  euler_update<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, f->x2, f->xn, f->vn, sp->h);
  MPRGP_HIP(hipGetLastError());
  MPRGP_HIP(hipStreamSynchronize(s));
  f->step_open = false;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_fem_advance(mprgp_fem *f, mprgp_qp *qp, const mprgp_step_params *sp, const mprgp_params *p,
                                 mprgp_step_stats *out) {
  MPRGP_TRY
  if (int rc = check_step(f, qp, sp)) return rc;
  if (!p) return fail(MPRGP_EBADINPUT, "NULL params");
  if (sp->newton_max_it < 1) return fail(MPRGP_EBADINPUT, "newton_max_it must be at least 1");
  if (!(sp->newton_eps >= 0.0)) return fail(MPRGP_EBADINPUT, "newton_eps must be >= 0");
  MPRGP_NOT_BUSY(qp);
  mprgp_ctx *ctx = f->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  mprgp_step_stats st;
  std::memset(&st, 0, sizeof(st));
  if (int rc = mprgp_fem_begin_step(f, qp, sp, &st.rows)) return rc;
  // forward (MprgpFemSolver.cpp:514-537): x = P(x1), solve, updatePos, until
  // max |x - x1| < eps or maxIter solves
  mprgp_params sp_solve = *p;
  sp_solve.project_first = 1;
  hipStream_t s = ctx->stream;
  const int N = f->N;
  const size_t n = 3 * (size_t)N;
  int worst = MPRGP_OK;
  for (int it = 0; it < sp->newton_max_it; ++it) {
    if (it > 0)
      if (int rc = assemble_dev(f, sp->material, f->xn, f->vn, f->x1, sp->gravity, sp->h, qp)) return rc;
    mprgp_stats ss;
    const int code = mprgp_qp_solve(qp, &sp_solve, &ss);
    if (code != MPRGP_OK && code != MPRGP_NOT_CONVERGED) return code;
    if (code != MPRGP_OK) worst = code;
    st.newton_iters = it + 1;
    st.iterations += ss.iterations;
    st.last = ss;
    MPRGP_HIP(launch_scatter_nodes(s, qp->x, f->x2, f->A->d_perm, N));
    if (int rc = ctx->scratch_reserve(sizeof(double) * (kMaxParts + 1))) return rc;
    MPRGP_HIP(launch_max_abs_diff(s, f->x2, f->x1, n, (double *)ctx->scratch));
    MPRGP_HIP(hipMemcpyAsync(&st.delta, ctx->scratch, sizeof(double), hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipStreamSynchronize(s));
    std::swap(f->x1, f->x2);  // x1 = the new iterate
    if (st.delta < sp->newton_eps) break;
  }
  // updateMesh from the last iterate (now in x1)
  euler_update<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, f->x1, f->xn, f->vn, sp->h);
  MPRGP_HIP(hipGetLastError());
  MPRGP_HIP(hipStreamSynchronize(s));
  f->step_open = false;
  st.code = worst;
  if (out) *out = st;
  return worst;
  MPRGP_CATCH
}

extern "C" void mprgp_fem_destroy(mprgp_fem *f) {
  MPRGP_TRY
  if (!f) return;
  mprgp_ctx *ctx = f->ctx;
  (void)hipSetDevice(ctx->device);
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->release(f->dev, f->dev_bytes);
  if (f->state) ctx->release(f->state, sizeof(double) * 4 * 3 * (size_t)f->N);
  mprgp_matrix_destroy(f->A);
  delete f;
  MPRGP_CATCH_VOID
}
