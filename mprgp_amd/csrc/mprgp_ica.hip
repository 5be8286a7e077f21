// mprgp_ica.hip -- the ICA comparator on the device: ICASolver
// (ICASolver.h:287-463; Otaduy et al., "Implicit Contact Handling for
// Deformable Objects", EG 2009), the solver the reference benchmarks MPRGP
// against (ICAFemSolver::forward, MprgpFemSolver.cpp:439-486), for
//     A x = b + J^T lambda,  0 <= lambda  _|_  J x - p >= 0.
//
// A = La + Da + Ua (3x3 diagonal blocks Da; La / Ua the blocks whose node
// comes before / after the row's node in the user's order).  Each iteration:
//     c      = p + J Da^-1 (LaUa x - b)        ica_rows(kOffDiag), J u
//     lambda = projected Gauss-Seidel on B = J Da^-1 J^T from the previous
//              lambda (host: B is m x m and the sweep is serial)
//     rhs    = J^T lambda - Ua x + b           ica_rows(kUpper)
//     x      = one forward block Gauss-Seidel sweep of (Da + La) x = rhs
//     |A x - b - J^T lambda| <= tol stops      ica_rows(kResidual), dot
//
// The sweep is sequential in the user's node order, and it stays so: nodes
// are grouped into dependency levels (a node's level is one more than the
// highest level among the nodes it couples to that come before it), and one
// workgroup runs the levels in order with a barrier between them.  Every
// node then sees exactly the new and old values the sequential sweep gives
// it, so the sweep is the reference's, not a multicolour variant with other
// iterates.  Each row adds its terms in the order of its blocks (the input's
// column order), singly rounded, as the oracle does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mprgp_host.h"

// no fused multiply-adds: each product and sum is rounded once, in the
// oracle's order
#pragma clang fp contract(off)

namespace mprgp {
namespace {

enum IcaRowMode { kOffDiag = 0, kUpper = 1, kResidual = 2 };

// 3x3 inverse (row-major) by the adjugate: inv(i,j) = cof(j,i) / det.
__device__ __forceinline__ void inv3(const double *m, double *inv) {
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

// Da^-1 per internal node: the row's block on its own node (padding slots
// also point at the row's node and add zeros).
__global__ void __launch_bounds__(256) ica_dinv(const int *sl_off, const int *col, const double *val, int n_nodes,
                                                double *dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int s = i / kSlice, l = i % kSlice;
  const int b0 = sl_off[s], W = (sl_off[s + 1] - b0) / kSlice;
  double D[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < W; ++k) {
    if (col[b0 + k * kSlice + l] != i) continue;
    const double *v = val + (size_t)b0 * 9 + (size_t)k * 9 * kSlice + l;
    for (int e = 0; e < 9; ++e) D[e] += v[e * kSlice];
  }
  inv3(D, dinv + 9 * (size_t)i);
}

// One pass over each row's blocks (internal node order, 3 values per node):
//   kOffDiag:  out = Da^-1 (sum over blocks off the node of A x  -  b)
//   kUpper:    out = (jtl - sum over blocks after the node of A x) + b
//   kResidual: out = (A x - b) - jtl
__global__ void __launch_bounds__(256) ica_rows(const int *sl_off, const int *col, const double *val,
                                                const int *perm, const double *x, const double *b,
                                                const double *jtl, const double *dinv, double *out, int n_nodes,
                                                int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int s = i / kSlice, l = i % kSlice;
  const int b0 = sl_off[s], W = (sl_off[s + 1] - b0) / kSlice;
  const int ui = perm[i];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int k = 0; k < W; ++k) {
    const int c = col[b0 + k * kSlice + l];
    const bool take = mode == kOffDiag ? c != i : (mode == kUpper ? perm[c] > ui : true);
    if (!take) continue;
    const double *v = val + (size_t)b0 * 9 + (size_t)k * 9 * kSlice + l;
    const double x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
    a0 += v[0 * kSlice] * x0;
    a0 += v[1 * kSlice] * x1;
    a0 += v[2 * kSlice] * x2;
    a1 += v[3 * kSlice] * x0;
    a1 += v[4 * kSlice] * x1;
    a1 += v[5 * kSlice] * x2;
    a2 += v[6 * kSlice] * x0;
    a2 += v[7 * kSlice] * x1;
    a2 += v[8 * kSlice] * x2;
  }
  const size_t r = 3 * (size_t)i;
  if (mode == kOffDiag) {
    const double t0 = a0 - b[r], t1 = a1 - b[r + 1], t2 = a2 - b[r + 2];
    const double *D = dinv + 9 * (size_t)i;
    out[r] = D[0] * t0 + D[1] * t1 + D[2] * t2;
    out[r + 1] = D[3] * t0 + D[4] * t1 + D[5] * t2;
    out[r + 2] = D[6] * t0 + D[7] * t1 + D[8] * t2;
  } else if (mode == kUpper) {
    out[r] = (jtl[r] - a0) + b[r];
    out[r + 1] = (jtl[r + 1] - a1) + b[r + 1];
    out[r + 2] = (jtl[r + 2] - a2) + b[r + 2];
  } else {
    out[r] = (a0 - b[r]) - jtl[r];
    out[r + 1] = (a1 - b[r + 1]) - jtl[r + 1];
    out[r + 2] = (a2 - b[r + 2]) - jtl[r + 2];
  }
}

// The lower blocks' values in sweep order, 9 per block (row-major), gathered
// once per call from their SELL slots: a slot's 9 values, for slice base b0
// and block k with slot = b0 + 64 k + l, sit at 9 (b0 + 64 k) + l + 64 e.
__global__ void __launch_bounds__(256) ica_lower_vals(const double *val, const int *lo_slot, int n_lo,
                                                      double *lo_val) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_lo) return;
  const long long slot = lo_slot[e];
  const double *v = val + (9 * slot - 8 * (slot % kSlice));
#pragma unroll
  for (int k = 0; k < 9; ++k) lo_val[9 * (size_t)e + k] = v[k * kSlice];
}

// The forward block Gauss-Seidel sweep (ICASolver.h:394-404) in one
// workgroup, level by level, one thread per node of a level.  A level's
// nodes depend only on nodes of earlier levels, whose x the barrier after
// each level publishes.  A node's lower blocks (its blocks on nodes before
// it) are listed in sweep order, lo_ptr[t] .. lo_ptr[t + 1] for the
// level-ordered node t, with their values copied in that order (lo_val) and
// their columns; the thread loads them kChunk at a time, every load of a
// chunk issued before the chunk's products, and subtracts v_ab x_b (each
// product rounded once: no contraction in this file) from rhs in block
// order, b = 0, 1, 2, as the sequential sweep does; then x = Da^-1 r.
constexpr int kChunk = 6;
__global__ void __launch_bounds__(512) ica_sweep(const double *lo_val, const int *lo_col, const int *lo_ptr,
                                                 const double *dinv, const double *rhs, double *x,
                                                 const int *lvl_ptr, const int *lvl_node, int n_levels) {
  for (int lv = 0; lv < n_levels; ++lv) {
    const int t1 = lvl_ptr[lv + 1];
    for (int t = lvl_ptr[lv] + (int)threadIdx.x; t < t1; t += (int)blockDim.x) {
      const int i = lvl_node[t], k0 = lo_ptr[t], k1 = lo_ptr[t + 1];
      const size_t r = 3 * (size_t)i;
      double r0 = rhs[r], r1 = rhs[r + 1], r2 = rhs[r + 2];
      const double *D = dinv + 9 * (size_t)i;
      for (int e0 = k0; e0 < k1; e0 += kChunk) {
        double v[kChunk][9], xb[kChunk][3];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          const int e = min(e0 + u, k1 - 1);  // past the node's last block: its last again, not used
          const int c = lo_col[e];
#pragma unroll
          for (int k = 0; k < 9; ++k) v[u][k] = lo_val[9 * (size_t)e + k];
#pragma unroll
          for (int k = 0; k < 3; ++k) xb[u][k] = x[3 * (size_t)c + k];
        }
        asm volatile("" ::: "memory");  // the chunk's loads stay ahead of its products
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
          if (e0 + u >= k1) break;
          r0 -= v[u][0] * xb[u][0];
          r0 -= v[u][1] * xb[u][1];
          r0 -= v[u][2] * xb[u][2];
          r1 -= v[u][3] * xb[u][0];
          r1 -= v[u][4] * xb[u][1];
          r1 -= v[u][5] * xb[u][2];
          r2 -= v[u][6] * xb[u][0];
          r2 -= v[u][7] * xb[u][1];
          r2 -= v[u][8] * xb[u][2];
        }
      }
      x[r] = D[0] * r0 + D[1] * r1 + D[2] * r2;
      x[r + 1] = D[3] * r0 + D[4] * r1 + D[5] * r2;
      x[r + 2] = D[6] * r0 + D[7] * r1 + D[8] * r2;
    }
    __syncthreads();
  }
}

constexpr int kSweepThreads = 512;

// ---------------------------------------------------------------- host side

// B = J Da^-1 J^T (CSR, sorted columns) from J over the user's DoFs and the
// inverse blocks in user node order.  Entry (i, k) adds, over the nodes the
// two rows share in ascending order, J_k(v) . (Dinv_v J_i(v)), each product
// left to right.
struct HostCsr {
  std::vector<int> ptr, idx;
  std::vector<double> val;
};

// A row of J as (node, 3-vector) pairs, nodes ascending.
struct RowNodes {
  std::vector<int> node;
  std::vector<double> v;  // 3 per node
};

std::vector<RowNodes> rows_by_node(int m, const int *j_ptr, const int *j_col, const double *j_val) {
  std::vector<RowNodes> rows(m);
  std::vector<std::pair<int, double>> e;
  for (int i = 0; i < m; ++i) {
    e.clear();
    for (int k = j_ptr[i]; k < j_ptr[i + 1]; ++k) e.emplace_back(j_col[k], j_val[k]);
    std::stable_sort(e.begin(), e.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    RowNodes &R = rows[i];
    for (const auto &pr : e) {
      const int v = pr.first / 3;
      if (R.node.empty() || R.node.back() != v) {
        R.node.push_back(v);
        R.v.insert(R.v.end(), {0.0, 0.0, 0.0});
      }
      R.v[R.v.size() - 3 + pr.first % 3] += pr.second;
    }
  }
  return rows;
}

HostCsr build_b(const std::vector<RowNodes> &rows, int n_nodes, const std::vector<double> &dinv_user) {
  const int m = (int)rows.size();
  // the rows touching each node, ascending, with their slot in that row
  std::vector<std::vector<std::pair<int, int>>> at(n_nodes);
  for (int i = 0; i < m; ++i)
    for (int q = 0; q < (int)rows[i].node.size(); ++q) at[rows[i].node[q]].emplace_back(i, q);
  HostCsr B;
  B.ptr.assign(m + 1, 0);
  std::vector<double> acc(m, 0.0);
  std::vector<char> seen(m, 0);
  std::vector<int> cols;
  for (int i = 0; i < m; ++i) {
    cols.clear();
    const RowNodes &Ri = rows[i];
    for (int q = 0; q < (int)Ri.node.size(); ++q) {
      const int v = Ri.node[q];
      const double *a = &Ri.v[3 * q];
      const double *D = &dinv_user[9 * (size_t)v];
      double w[3];
      for (int r = 0; r < 3; ++r) w[r] = D[3 * r] * a[0] + D[3 * r + 1] * a[1] + D[3 * r + 2] * a[2];
      for (const auto &kq : at[v]) {
        const double *jk = &rows[kq.first].v[3 * kq.second];
        const double d = jk[0] * w[0] + jk[1] * w[1] + jk[2] * w[2];
        if (!seen[kq.first]) {
          seen[kq.first] = 1;
          cols.push_back(kq.first);
          acc[kq.first] = d;
        } else {
          acc[kq.first] += d;
        }
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int k : cols) {
      B.idx.push_back(k);
      B.val.push_back(acc[k]);
      seen[k] = 0;
    }
    B.ptr[i + 1] = (int)B.idx.size();
  }
  return B;
}

// GaussSeidel<true>::solve (ICASolver.h:68-111) on B from lam: rhs = c - Ub lam
// (old lam), then row by row rhs_r -= sum_{k<r} B_rk lam_k (new lam),
// lam_r = max(0, rhs_r / B_rr); stop when |lam - lam_prev| <= tol.  Returns
// the sweeps run.
int pgs(const HostCsr &B, const std::vector<double> &invd, const std::vector<double> &c, std::vector<double> &lam,
        int max_it, double tol, std::vector<double> &rhs, std::vector<double> &pre) {
  const int m = (int)c.size();
  for (int it = 0; it < max_it; ++it) {
    pre = lam;
    for (int r = 0; r < m; ++r) {
      double u = 0.0;
      for (int k = B.ptr[r]; k < B.ptr[r + 1]; ++k)
        if (B.idx[k] > r) u += B.val[k] * lam[B.idx[k]];
      rhs[r] = c[r] - u;
    }
    for (int r = 0; r < m; ++r) {
      for (int k = B.ptr[r]; k < B.ptr[r + 1]; ++k)
        if (B.idx[k] < r) rhs[r] -= lam[B.idx[k]] * B.val[k];
      const double v = rhs[r] * invd[r];
      lam[r] = v > 0 ? v : 0;
    }
    double s = 0.0;
    for (int r = 0; r < m; ++r) s += (lam[r] - pre[r]) * (lam[r] - pre[r]);
    if (std::sqrt(s) <= tol) return it + 1;
  }
  return max_it;
}

}  // namespace
}  // namespace mprgp

using namespace mprgp;

extern "C" int mprgp_ica_solve(mprgp_qp *q, int m, const int *j_ptr, const int *j_col, const double *j_val,
                               const double *p, double tol, int max_it, int gs_max_it, double gs_tol,
                               double *lambda, int *pgs_sweeps, mprgp_stats *stats) {
  MPRGP_TRY
  if (!q || m < 0 || (m > 0 && (!j_ptr || !j_col || !j_val || !p))) return fail(MPRGP_EBADINPUT, "bad ICA arguments");
  if (!(tol == tol) || max_it < 0 || gs_max_it < 0 || !(gs_tol == gs_tol))
    return fail(MPRGP_EBADINPUT, "bad ICA tolerances or iteration limits");
  if (!q->has_rhs || !q->has_x) return fail(MPRGP_ESTATE, "rhs and x must be set");
  MPRGP_NOT_BUSY(q);
  mprgp_ctx *ctx = q->ctx;
  const mprgp_matrix *A = q->A;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  const int nn = q->n, n = 3 * nn;
  RowsInternal Ri;
  if (int rc = rows_to_internal(A, m, j_ptr, j_col, j_val, p, Ri)) return rc;
  const int nnz = m > 0 ? j_ptr[m] : 0;
  hipStream_t s = ctx->stream;

  // dependency levels of the sweep from the matrix structure (user order)
  std::vector<int> col((size_t)A->slots);
  MPRGP_HIP(hipMemcpyAsync(col.data(), A->d_col, sizeof(int) * col.size(), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  std::vector<int> level(nn, 0);
  int n_levels = nn > 0 ? 1 : 0;
  for (int u = 0; u < nn; ++u) {
    const int i = A->iperm[u], sl = i / kSlice, l = i % kSlice;
    const int b0 = A->sl_off_h[sl], W = (A->sl_off_h[sl + 1] - b0) / kSlice;
    int lv = 0;
    for (int k = 0; k < W; ++k) {
      const int c = col[(size_t)b0 + (size_t)k * kSlice + l];
      const int uc = A->perm[c];
      if (uc < u) lv = std::max(lv, level[uc] + 1);
    }
    level[u] = lv;
    n_levels = std::max(n_levels, lv + 1);
  }
  std::vector<int> lvl_ptr(n_levels + 1, 0), lvl_node(nn);
  for (int u = 0; u < nn; ++u) lvl_ptr[level[u] + 1]++;
  for (int lv = 0; lv < n_levels; ++lv) lvl_ptr[lv + 1] += lvl_ptr[lv];
  {
    std::vector<int> fill(lvl_ptr.begin(), lvl_ptr.end() - 1);
    for (int u = 0; u < nn; ++u) lvl_node[fill[level[u]]++] = A->iperm[u];
  }
  // each level-ordered node's lower blocks (nodes before it), in block
  // order: SELL slot and column
  std::vector<int> lo_ptr(nn + 1, 0), lo_slot, lo_col;
  for (int t = 0; t < nn; ++t) {
    const int i = lvl_node[t], u = A->perm[i], sl = i / kSlice, l = i % kSlice;
    const int b0 = A->sl_off_h[sl], W = (A->sl_off_h[sl + 1] - b0) / kSlice;
    for (int k = 0; k < W; ++k) {
      const int slot = b0 + k * kSlice + l, c = col[(size_t)slot];
      if (A->perm[c] < u) {
        lo_slot.push_back(slot);
        lo_col.push_back(c);
      }
    }
    lo_ptr[t + 1] = (int)lo_slot.size();
  }
  const size_t n_lo = lo_slot.size();

  // device block: Da^-1, levels, the lower blocks and their products, J and
  // J^T, J u, lambda, dot partials
  const size_t parts = ((size_t)n + 255) / 256 + 1;
  const size_t bytes = sizeof(double) * (9 * (size_t)nn + 9 * n_lo + 2 * (size_t)nnz + 2 * (size_t)m + parts + 8) +
                       sizeof(int) * ((size_t)n_levels + 1 + (size_t)nn + (size_t)nn + 1 + 2 * n_lo + (size_t)m +
                                      1 + 2 * (size_t)nnz + (size_t)n + 1) +
                       1024;
  void *blk = nullptr;
  if (int rc = ctx->alloc(&blk, bytes)) return rc;
  struct Release {  // on every exit: the stream's work on the block ends first
    mprgp_ctx *c;
    void *p;
    size_t b;
    ~Release() {
      (void)hipStreamSynchronize(c->stream);
      c->release(p, b);
    }
  } release_{ctx, blk, bytes};
  char *b8 = (char *)blk;
  auto carve = [&](size_t sz) {
    uintptr_t u = ((uintptr_t)b8 + 7) & ~(uintptr_t)7;
    b8 = (char *)(u + sz);
    return (void *)u;
  };
  double *d_dinv = (double *)carve(sizeof(double) * 9 * (size_t)nn);
  double *d_jv = (double *)carve(sizeof(double) * nnz);
  double *d_tv = (double *)carve(sizeof(double) * nnz);
  double *d_jx = (double *)carve(sizeof(double) * (m + 1));
  double *d_lam = (double *)carve(sizeof(double) * (m + 1));
  double *d_part = (double *)carve(sizeof(double) * parts);
  double *d_res = (double *)carve(sizeof(double) * 8);
  int *d_lvl_ptr = (int *)carve(sizeof(int) * (n_levels + 1));
  int *d_lvl_node = (int *)carve(sizeof(int) * (size_t)nn);
  double *d_lo_val = (double *)carve(sizeof(double) * 9 * n_lo);
  int *d_lo_ptr = (int *)carve(sizeof(int) * ((size_t)nn + 1));
  int *d_lo_slot = (int *)carve(sizeof(int) * n_lo);
  int *d_lo_col = (int *)carve(sizeof(int) * n_lo);
  int *d_jp = (int *)carve(sizeof(int) * (m + 1));
  int *d_jc = (int *)carve(sizeof(int) * nnz);
  int *d_tp = (int *)carve(sizeof(int) * (n + 1));
  int *d_tc = (int *)carve(sizeof(int) * nnz);
  if (m > 0) {
    MPRGP_HIP(hipMemcpyAsync(d_jp, j_ptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(d_jc, Ri.jc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(d_jv, j_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(d_tc, Ri.tc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(d_tv, Ri.tv.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  }
  MPRGP_HIP(hipMemcpyAsync(d_tp, Ri.tptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(d_lvl_ptr, lvl_ptr.data(), sizeof(int) * (n_levels + 1), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(d_lvl_node, lvl_node.data(), sizeof(int) * (size_t)nn, hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(d_lo_ptr, lo_ptr.data(), sizeof(int) * ((size_t)nn + 1), hipMemcpyHostToDevice, s));
  if (n_lo > 0) {
    MPRGP_HIP(hipMemcpyAsync(d_lo_slot, lo_slot.data(), sizeof(int) * n_lo, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(d_lo_col, lo_col.data(), sizeof(int) * n_lo, hipMemcpyHostToDevice, s));
  }
  const int blocks = (nn + 255) / 256;
  if (nn > 0) {
    ica_dinv<<<blocks, 256, 0, s>>>(A->d_sl_off, A->d_col, A->d_val, nn, d_dinv);
    MPRGP_HIP(hipGetLastError());
  }
  if (n_lo > 0) {
    ica_lower_vals<<<(unsigned)((n_lo + 255) / 256), 256, 0, s>>>(A->d_val, d_lo_slot, (int)n_lo, d_lo_val);
    MPRGP_HIP(hipGetLastError());
  }
  // B = J Da^-1 J^T on the host, from the inverse blocks in user order
  std::vector<RowNodes> rows;
  HostCsr B;
  std::vector<double> invd(m, 0.0), c(m), lam(m, 0.0), rhs(m), pre(m), jx(m);
  if (m > 0) {
    std::vector<double> dinv((size_t)9 * nn), dinv_user((size_t)9 * nn);
    MPRGP_HIP(hipMemcpyAsync(dinv.data(), d_dinv, sizeof(double) * dinv.size(), hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipStreamSynchronize(s));
    for (int u = 0; u < nn; ++u)
      std::copy(&dinv[9 * (size_t)A->iperm[u]], &dinv[9 * (size_t)A->iperm[u]] + 9, &dinv_user[9 * (size_t)u]);
    rows = rows_by_node(m, j_ptr, j_col, j_val);
    B = build_b(rows, nn, dinv_user);
    for (int r = 0; r < m; ++r)
      for (int k = B.ptr[r]; k < B.ptr[r + 1]; ++k)
        if (B.idx[k] == r) invd[r] = 1.0 / B.val[k];
  }
  double *u = q->z, *jtl = q->phi, *rh = q->beta, *e = q->ap;
  const double *x_c = q->x;
  MPRGP_HIP(hipMemsetAsync(jtl, 0, sizeof(double) * n, s));
  MPRGP_HIP(hipMemsetAsync(d_res, 0, sizeof(double) * 8, s));
  MPRGP_HIP(hipEventRecord(ctx->t0, s));
  int it = 0, sweeps = 0;
  bool met = false;
  double residual = -1.0;
  for (it = 0; it < max_it; ++it) {
    if (m > 0) {
      ica_rows<<<blocks, 256, 0, s>>>(A->d_sl_off, A->d_col, A->d_val, A->d_perm, x_c, q->b, jtl, d_dinv, u, nn,
                                      kOffDiag);
      MPRGP_HIP(hipGetLastError());
      MPRGP_HIP(launch_csr_mul(s, m, d_jp, d_jc, d_jv, u, d_jx));
      MPRGP_HIP(hipMemcpyAsync(jx.data(), d_jx, sizeof(double) * m, hipMemcpyDeviceToHost, s));
      MPRGP_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < m; ++i) c[i] = p[i] + jx[i];
      sweeps += pgs(B, invd, c, lam, gs_max_it, gs_tol, rhs, pre);
      MPRGP_HIP(hipMemcpyAsync(d_lam, lam.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
      MPRGP_HIP(launch_csr_mul(s, n, d_tp, d_tc, d_tv, d_lam, jtl));
    }
    if (nn > 0) {
      ica_rows<<<blocks, 256, 0, s>>>(A->d_sl_off, A->d_col, A->d_val, A->d_perm, x_c, q->b, jtl, d_dinv, rh, nn,
                                      kUpper);
      MPRGP_HIP(hipGetLastError());
      ica_sweep<<<1, kSweepThreads, 0, s>>>(d_lo_val, d_lo_col, d_lo_ptr, d_dinv, rh, q->x, d_lvl_ptr,
                                            d_lvl_node, n_levels);
      MPRGP_HIP(hipGetLastError());
      ica_rows<<<blocks, 256, 0, s>>>(A->d_sl_off, A->d_col, A->d_val, A->d_perm, x_c, q->b, jtl, d_dinv, e, nn,
                                      kResidual);
      MPRGP_HIP(hipGetLastError());
    }
    MPRGP_HIP(launch_dot(s, e, e, (size_t)n, d_part, d_res));
    double r2 = 0.0;
    MPRGP_HIP(hipMemcpyAsync(&r2, d_res, sizeof(double), hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipStreamSynchronize(s));  // also completes the lambda upload before the next PGS
    residual = std::sqrt(r2);
    if (residual <= tol) {
      met = true;
      break;
    }
  }
  MPRGP_HIP(hipEventRecord(ctx->t1, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  MPRGP_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  if (lambda)
    for (int i = 0; i < m; ++i) lambda[i] = lam[i];
  if (pgs_sweeps) *pgs_sweeps = sweeps;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->code = met ? 0 : -1;
    stats->iterations = it;
    stats->residual = residual;
    stats->elapsed_ms = ms;
    stats->error_bound = -1.0;
  }
  if (!std::isfinite(residual) && max_it > 0) return fail(MPRGP_ENAN, "ICA residual is not finite");
  return met ? MPRGP_OK : MPRGP_NOT_CONVERGED;
  MPRGP_CATCH
}
