// mprgp_general.cpp -- the coupled-constraint solve: MPRGPGeneralCon<double>::
// solve (MPRGPSolver.h:594-626) with GeneralConProjector
// (MPRGPProjection.h:400-520) and DiagonalInFacePreconSolver
// (MPRGPPrecondition.h:36-69), over MPRGPMonotonic::solve (MPRGPSolver.h:
// 148-316, 412-454).
//
// Rows J x >= c may span several nodes (the self-contact rows of
// LinearConCollider.cpp:186-217: a vertex against a triangle, 12 nonzeros),
// so the projection is a QP in the rows' multipliers rather than a per-node
// formula.  The n-sized work stays on the device: SpMVs with A (the QP's SELL
// matrix), J v and J^T lambda (CSR, one thread per row), vector updates and
// dot products.  The m-sized work runs on the host: the dense QP
// min 1/2 l'(J J^T) l - h'l, l >= 0 of each projection and of BETA, and the
// Cholesky factor of the active rows' J_a J_a^T for PHI -- where the
// reference calls QuadProg++ and Eigen's SimplicialLLT.  J J^T is formed
// once per solve.  Each of these decisions reads a few device scalars or an
// m-vector back, so a solve synchronises the stream a few times per
// iteration: this path is for coupled rows, which the simulator's decoupled
// solve (the batched kernels) does not have.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mprgp_host.h"

using namespace mprgp;

namespace {

constexpr int kMaxGeneralRows = 1 << 20;
constexpr int kMaxComponentRows = 2048;  // a component's block of J J^T is dense on the host
constexpr size_t kMaxCoupledPairs = (size_t)1 << 26;  // J J^T entries formed (16 bytes each on the host)

// Dense Cholesky of a k x k row-major SPD block in place (lower factor);
// false if a pivot is not positive.
bool chol_factor(std::vector<double> &L, int k) {
  for (int j = 0; j < k; ++j) {
    double d = L[(size_t)j * k + j];
    for (int p = 0; p < j; ++p) d -= L[(size_t)j * k + p] * L[(size_t)j * k + p];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    L[(size_t)j * k + j] = d;
    for (int i = j + 1; i < k; ++i) {
      double v = L[(size_t)i * k + j];
      for (int p = 0; p < j; ++p) v -= L[(size_t)i * k + p] * L[(size_t)j * k + p];
      L[(size_t)i * k + j] = v / d;
    }
  }
  return true;
}

void chol_solve(const std::vector<double> &L, int k, double *b) {
  for (int i = 0; i < k; ++i) {
    double v = b[i];
    for (int p = 0; p < i; ++p) v -= L[(size_t)i * k + p] * b[p];
    b[i] = v / L[(size_t)i * k + i];
  }
  for (int i = k - 1; i >= 0; --i) {
    double v = b[i];
    for (int p = i + 1; p < k; ++p) v -= L[(size_t)p * k + i] * b[p];
    b[i] = v / L[(size_t)i * k + i];
  }
}

// lam = argmin 1/2 l'Gl - h'l subject to l >= 0, G (k x k, row-major) SPD.
// Primal active set on the free set F: free the bound variable with the
// largest positive residual (h - G l)_i, solve G_FF s = h_F by Cholesky,
// and while some s_i <= 0 step towards s up to the first zero and bound it.
// Finite for SPD G; false if a free set's G_FF is not positive definite.
// tol: the residual a bound variable needs to be freed (qp_tol below).
// The freeing threshold: 1e-14 max |h| (1e-14 if h = 0).
double qp_tol(const double *h, int k) {
  double hs = 0.0;
  for (int i = 0; i < k; ++i) hs = std::max(hs, std::fabs(h[i]));
  return 1e-14 * (hs > 0.0 ? hs : 1.0);
}

bool qp_nonneg(int k, const std::vector<double> &G, const std::vector<double> &h, std::vector<double> &lam,
               double tol) {
  lam.assign(k, 0.0);
  if (k == 0) return true;
  std::vector<char> free_(k, 0);
  std::vector<int> idx;
  std::vector<double> L, s;
  for (int outer = 0; outer < 10 * k + 10; ++outer) {
    int best = -1;
    double wbest = tol;
    for (int i = 0; i < k; ++i) {
      double r = h[i];
      for (int j = 0; j < k; ++j) r -= G[(size_t)i * k + j] * lam[j];
      if (!free_[i] && r > wbest) {
        wbest = r;
        best = i;
      }
    }
    if (best < 0) return true;
    free_[best] = 1;
    for (int inner = 0; inner <= k; ++inner) {
      idx.clear();
      for (int i = 0; i < k; ++i)
        if (free_[i]) idx.push_back(i);
      const int f = (int)idx.size();
      L.assign((size_t)f * f, 0.0);
      s.assign(f, 0.0);
      for (int a = 0; a < f; ++a) {
        for (int b = 0; b < f; ++b) L[(size_t)a * f + b] = G[(size_t)idx[a] * k + idx[b]];
        s[a] = h[idx[a]];
      }
      if (!chol_factor(L, f)) return false;
      chol_solve(L, f, s.data());
      double step = 1.0;
      int hit = -1;
      for (int a = 0; a < f; ++a)
        if (s[a] <= 0.0) {
          const double t = lam[idx[a]] / (lam[idx[a]] - s[a]);
          if (t < step) {
            step = t;
            hit = a;
          }
        }
      if (hit < 0) {
        std::fill(lam.begin(), lam.end(), 0.0);
        for (int a = 0; a < f; ++a) lam[idx[a]] = s[a];
        break;
      }
      for (int a = 0; a < f; ++a) lam[idx[a]] += step * (s[a] - lam[idx[a]]);
      free_[idx[hit]] = 0;
      lam[idx[hit]] = 0.0;
      for (int a = 0; a < f; ++a)
        if (lam[idx[a]] <= 0.0) {
          free_[idx[a]] = 0;
          lam[idx[a]] = 0.0;
        }
    }
  }
  return false;
}

// One coupled-row solve on a QP: device vectors are the QP's own buffers
// (internal node order); J and J^T live in one call-owned device block.
struct General {
  mprgp_qp *q = nullptr;
  mprgp_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  int n = 0, m = 0;
  // device
  void *blk = nullptr;
  size_t blk_bytes = 0;
  int *jp = nullptr, *jc = nullptr, *tp = nullptr, *tc = nullptr;
  double *jv = nullptr, *tv = nullptr, *lam = nullptr, *part = nullptr, *res = nullptr;
  // tl holds J^T lambda inside jt_apply only, so a caller may pass any of
  // the other vectors (t included) as its input
  double *x, *g, *p, *z, *phi, *beta, *y, *gp, *t, *t2, *t3, *tl, *pv, *pw;
  const double *b = nullptr, *dinv = nullptr;
  // host
  std::vector<double> c;
  std::vector<char> face;
  std::vector<int> act;
  // J J^T is block diagonal over the connected components of its pattern
  // (rows that share a node): the projection's QP and the active rows'
  // factor split into one small problem per component.  comps: each
  // component's rows, ascending; comp_of/local_of: a row's component and
  // place in it; gc: each component's block of J J^T, dense, row-major;
  // ablk/aL: the active rows of each component with nonempty ones, and the
  // Cholesky factor of their J_a J_a^T block.
  std::vector<std::vector<int>> comps, ablk;
  std::vector<int> comp_of, local_of;
  std::vector<std::vector<double>> gc, aL;
  double jjt(int i, int j) const {
    const int c0 = comp_of[i];
    if (comp_of[j] != c0) return 0.0;
    return gc[c0][(size_t)local_of[i] * comps[c0].size() + local_of[j]];
  }
  bool qp_fail = false;
  int y_size = 0;

  // J J^T on the pairs of rows that share a column (J^T's pattern: column
  // -> rows), each entry formed as a dense pass forms it: row i scattered
  // into a dense vector, row j's terms against it in stored order.  Rows are
  // joined where an entry is nonzero (other pairs are exactly 0), and each
  // connected component keeps its dense block.
  int build_jjt(const int *j_ptr, const double *j_val, const std::vector<int> &jc, const std::vector<int> &tptr,
                const std::vector<int> &tc) {
    std::vector<std::vector<std::pair<int, double>>> nz(m);
    std::vector<double> row(n, 0.0);
    std::vector<int> mark(m, -1), cand, parent(m);
    size_t pairs = 0;
    for (int i = 0; i < m; ++i) parent[i] = i;
    auto find = [&](int i) {
      while (parent[i] != i) i = parent[i] = parent[parent[i]];
      return i;
    };
    for (int i = 0; i < m; ++i) {
      cand.clear();
      for (int k = j_ptr[i]; k < j_ptr[i + 1]; ++k) {
        row[jc[k]] += j_val[k];
        for (int t = tptr[jc[k]]; t < tptr[jc[k] + 1]; ++t)
          if (mark[tc[t]] != i) {
            mark[tc[t]] = i;
            cand.push_back(tc[t]);
          }
      }
      // rows that share a column with row i all join its component unless
      // their products vanish; bound the work and memory up front
      if (cand.size() > (size_t)kMaxComponentRows)
        return fail(MPRGP_EBADINPUT, "row %d shares a column with %zu rows; at most %d in one component", i,
                    cand.size(), kMaxComponentRows);
      pairs += cand.size();
      if (pairs > kMaxCoupledPairs)
        return fail(MPRGP_EBADINPUT, "more than %zu pairs of rows share a column", kMaxCoupledPairs);
      std::sort(cand.begin(), cand.end());
      for (int j : cand) {
        double v = 0.0;
        for (int k = j_ptr[j]; k < j_ptr[j + 1]; ++k) v += j_val[k] * row[jc[k]];
        nz[i].emplace_back(j, v);
        if (v != 0.0 && j != i) {
          const int a = find(i), b2 = find(j);
          if (a != b2) parent[std::max(a, b2)] = std::min(a, b2);
        }
      }
      for (int k = j_ptr[i]; k < j_ptr[i + 1]; ++k) row[jc[k]] = 0.0;
    }
    std::vector<int> slot(m, -1);
    comps.clear();
    comp_of.assign(m, 0);
    local_of.assign(m, 0);
    for (int i = 0; i < m; ++i) {
      const int r = find(i);
      if (slot[r] < 0) {
        slot[r] = (int)comps.size();
        comps.emplace_back();
      }
      comp_of[i] = slot[r];
      local_of[i] = (int)comps[slot[r]].size();
      comps[slot[r]].push_back(i);
    }
    gc.assign(comps.size(), {});
    for (size_t c0 = 0; c0 < comps.size(); ++c0) {
      const size_t k = comps[c0].size();
      if (k > (size_t)kMaxComponentRows)
        return fail(MPRGP_EBADINPUT, "%zu rows coupled through shared nodes; at most %d in one component", k,
                    kMaxComponentRows);
      gc[c0].assign(k * k, 0.0);
      for (int i : comps[c0])
        for (const auto &e : nz[i])
          if (comp_of[e.first] == (int)c0) gc[c0][(size_t)local_of[i] * k + local_of[e.first]] = e.second;
    }
    return MPRGP_OK;
  }
  // l[rows] = argmin 1/2 l'Gl - h'l, l >= 0, over J J^T's block of `rows`
  // (h indexed by row), freeing threshold tol
  void qp_rows(const std::vector<int> &rows, const double *h, double tol, std::vector<double> &l) {
    const int k = (int)rows.size();
    std::vector<double> Gk, hk(k), lk;
    for (int a = 0; a < k; ++a) hk[a] = h[rows[a]];
    if (k == (int)comps[comp_of[rows[0]]].size()) {  // a whole component: its block
      Gk = gc[comp_of[rows[0]]];
    } else {
      Gk.resize((size_t)k * k);
      for (int a = 0; a < k; ++a)
        for (int bb = 0; bb < k; ++bb) Gk[(size_t)a * k + bb] = jjt(rows[a], rows[bb]);
    }
    if (!qp_nonneg(k, Gk, hk, lk, tol)) qp_fail = true;
    for (int a = 0; a < k; ++a) l[rows[a]] = lk.empty() ? 0.0 : lk[a];
  }

  // Pinned host memory: the multipliers' staging (two slots of m + 1) and
  // the copy of the device readback area.  Every fill of the staging
  // follows a readback that waited for the stream after the previous upload
  // was queued, so no upload still reads the slots it overwrites.
  double *h_lam = nullptr, *h_rb = nullptr;
  // Device readback area: three J v slots of m + 1, then 16 dot results.
  double *rb = nullptr;
  double *jslot(int k) const { return rb + (size_t)k * (m + 1); }
  const double *hj(int k) const { return h_rb + (size_t)k * (m + 1); }
  const double *hd() const { return h_rb + 3 * ((size_t)m + 1); }
  // host copies of J x and J y, valid while the vector is unchanged
  std::vector<double> jx_h, jy_h;
  bool jx_ok = false, jy_ok = false;

  ~General() {
    if (h_lam) {
      (void)hipStreamSynchronize(s);  // an upload from it may still be queued
      (void)hipHostFree(h_lam);
    }
    if (blk) ctx->release(blk, blk_bytes);
  }

  // queue J v into readback slot k
  int jlaunch(const double *v, int k) {
    MPRGP_HIP(launch_csr_mul(s, m, jp, jc, jv, v, jslot(k)));
    return MPRGP_OK;
  }
  // queue the dot products of the pairs into results first, first + 1, ...
  // (one kernel pair)
  int dgroup(std::initializer_list<std::pair<const double *, const double *>> pairs, int first) {
    if (pairs.size() > (size_t)kMaxDots) return fail(MPRGP_EHIP, "%zu dot products in one group", pairs.size());
    const double *a[kMaxDots], *b2[kMaxDots];
    int k = 0;
    for (const auto &pr : pairs) {
      a[k] = pr.first;
      b2[k] = pr.second;
      ++k;
    }
    MPRGP_HIP(launch_dots(s, k, a, b2, (size_t)n, part, res + first));
    return MPRGP_OK;
  }
  // J v slots [0, slots) and dot results [0, k) to the host, one wait
  int readback(int slots, int k) {
    if (slots > 0)
      MPRGP_HIP(hipMemcpyAsync(h_rb, rb, sizeof(double) * (size_t)slots * (m + 1), hipMemcpyDeviceToHost, s));
    if (k > 0)
      MPRGP_HIP(hipMemcpyAsync(h_rb + 3 * ((size_t)m + 1), res, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipStreamSynchronize(s));
    return MPRGP_OK;
  }
  // J v[j] into readback slot[j], j < k, in one launch
  int jlaunch_n(int k, const double *const *v, const int *slot) {
    double *o[kMaxCsrVecs];
    for (int j = 0; j < k; ++j) o[j] = jslot(slot[j]);
    MPRGP_HIP(launch_csr_mul_n(s, m, jp, jc, jv, k, v, o));
    return MPRGP_OK;
  }
  // J v to the host (slot 0)
  int jmul(const double *v) {
    if (int rc = jlaunch(v, 0)) return rc;
    return readback(1, 0);
  }
  // multipliers l into staging slot k
  void stage(int k, const std::vector<double> &l) { std::copy(l.begin(), l.end(), h_lam + (size_t)k * (m + 1)); }
  // upload staging slots [0, k), no wait of its own
  int upload(int k) {
    MPRGP_HIP(hipMemcpyAsync(lam, h_lam, sizeof(double) * (size_t)k * (m + 1), hipMemcpyHostToDevice, s));
    return MPRGP_OK;
  }
  // out = v + sign J^T l for the multipliers uploaded into slot k
  int jt_apply(int k, const double *v, double sign, double *out) {
    MPRGP_HIP(launch_csr_mul(s, n, tp, tc, tv, lam + (size_t)k * (m + 1), tl));
    MPRGP_HIP(launch_axpby(s, n, 1.0, v, sign, tl, out));
    return MPRGP_OK;
  }
  // dot products of the pairs, all read back with one synchronisation
  int dots(std::initializer_list<std::pair<const double *, const double *>> pairs, double *out) {
    const int k = (int)pairs.size();
    if (int rc = dgroup(pairs, 0)) return rc;
    if (int rc = readback(0, k)) return rc;
    std::copy(hd(), hd() + k, out);
    return MPRGP_OK;
  }
  int spmv(const double *v, double *out) {
    const mprgp_matrix *A = q->A;
    MPRGP_HIP(launch_spmv_plain(s, A->d_sl_off, A->d_col, A->d_val, v, out, A->n_nodes, 0));
    return MPRGP_OK;
  }
  int grad(const double *v) {  // g = A v - b
    if (int rc = spmv(v, g)) return rc;
    MPRGP_HIP(launch_axpby(s, n, 1.0, g, -1.0, b, g));
    return MPRGP_OK;
  }
  // queue 1/2 v'Av and v'b (FixedSparseMatrix::funcValue) into dot results
  // i, i + 1: f = 0.5 d[i] - d[i + 1]
  int func_launch(const double *v, int i) {
    if (int rc = spmv(v, t3)) return rc;
    return dgroup({{v, t3}, {v, b}}, i);
  }

  // GeneralConProjector::project (:412-431), the multipliers: one QP per
  // component of J J^T, h = c - J v, with the freeing threshold of the
  // whole QP
  void proj_lam(const double *jv_h, std::vector<double> &l) {
    std::vector<double> h(m);
    for (int i = 0; i < m; ++i) h[i] = c[i] - jv_h[i];
    const double tol = qp_tol(h.data(), m);
    l.assign(m, 0.0);
    for (const auto &rows : comps) qp_rows(rows, h.data(), tol, l);
  }
  // out = P(v); jv_h: J v on the host when already known
  int project(const double *v, double *out, const double *jv_h = nullptr) {
    if (m == 0) {
      if (out != v) MPRGP_HIP(hipMemcpyAsync(out, v, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      return MPRGP_OK;
    }
    if (!jv_h) {
      if (int rc = jmul(v)) return rc;
      jv_h = hj(0);
    }
    std::vector<double> l;
    proj_lam(jv_h, l);
    stage(0, l);
    if (int rc = upload(1)) return rc;
    return jt_apply(0, v, 1.0, out);
  }
  // DECIDE_FACE + prepareActiveConMatrix (:485-520) at v, then PHI
  // (:432-451) at the current g: J v and J g come back together.  With v = x
  // the host keeps J x.
  int face_phi(const double *v) {
    act.clear();
    ablk.clear();
    aL.clear();
    if (m == 0) {
      MPRGP_HIP(hipMemcpyAsync(phi, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      return MPRGP_OK;
    }
    const double *vs[2] = {v, g};
    const int sl[2] = {0, 1};
    if (int rc = jlaunch_n(2, vs, sl)) return rc;
    if (int rc = readback(2, 0)) return rc;
    const double *jv_h = hj(0), *jg = hj(1);
    if (v == x) {
      jx_h.assign(jv_h, jv_h + m);
      jx_ok = true;
    }
    for (int i = 0; i < m; ++i) {
      face[i] = std::fabs(jv_h[i] - c[i]) < kEps ? 1 : 0;
      if (face[i]) act.push_back(i);
    }
    for (const auto &rows : comps) {
      std::vector<int> a;
      for (int i : rows)
        if (face[i]) a.push_back(i);
      if (a.empty()) continue;
      const int k = (int)a.size();
      std::vector<double> L((size_t)k * k);
      for (int x0 = 0; x0 < k; ++x0)
        for (int x1 = 0; x1 < k; ++x1) L[(size_t)x0 * k + x1] = jjt(a[x0], a[x1]);
      if (!chol_factor(L, k)) qp_fail = true;
      ablk.push_back(std::move(a));
      aL.push_back(std::move(L));
    }
    if (act.empty()) {
      MPRGP_HIP(hipMemcpyAsync(phi, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      return MPRGP_OK;
    }
    return phi_from(jg);
  }
  // PHI (:432-451) from J g on the host: phi = g + J_a^T l, l solving
  // (J_a J_a^T) l = -J_a g per component
  int phi_from(const double *jg) {
    std::vector<double> l(m, 0.0);
    for (size_t blk_i = 0; blk_i < ablk.size(); ++blk_i) {
      const std::vector<int> &a = ablk[blk_i];
      const int k = (int)a.size();
      std::vector<double> la(k);
      for (int x0 = 0; x0 < k; ++x0) la[x0] = -jg[a[x0]];
      chol_solve(aL[blk_i], k, la.data());
      for (int x0 = 0; x0 < k; ++x0) l[a[x0]] = la[x0];
    }
    stage(0, l);
    if (int rc = upload(1)) return rc;
    return jt_apply(0, g, 1.0, phi);
  }
  // PHI at the current g and face
  int compute_phi() {
    if (act.empty()) {
      MPRGP_HIP(hipMemcpyAsync(phi, g, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      return MPRGP_OK;
    }
    if (int rc = jmul(g)) return rc;
    return phi_from(hj(0));
  }
  bool feasible_h(const double *jv_h) const {  // isFeasible (:298-310) from J v
    for (int i = 0; i < m; ++i)
      if (jv_h[i] < c[i] - kEps) return false;
    return true;
  }

  // computeGradients without PHI (:163-173) and proportional (:174-181) at
  // v, with phi current.  First half: beta0 = g - phi and t2 = v - ab phi,
  // J beta0 (slot 0, for BETA :452-473) and J t2 (slot 1, for PHITPHI's
  // projection :474-484) queued, and J y (slot 2) when want_jy; the caller
  // adds its own dot products from result 2 on and reads everything back
  // with head_wait.
  int head_launch(const double *v, double ab, bool want_jy) {
    MPRGP_HIP(launch_axpby(s, n, 1.0, g, -1.0, phi, beta));
    MPRGP_HIP(launch_axpby(s, n, 1.0, v, -ab, phi, t2));
    head_slots = 0;
    if (m == 0) return MPRGP_OK;
    const double *vs[3];
    int sl[3], k = 0;
    if (!act.empty()) {
      vs[k] = beta;
      sl[k++] = 0;
    }
    vs[k] = t2;
    sl[k++] = 1;
    head_slots = 2;
    if (want_jy) {
      vs[k] = y;
      sl[k++] = 2;
      head_slots = 3;
    }
    return jlaunch_n(k, vs, sl);
  }
  int head_slots = 0;
  int head_wait(int k) {
    if (int rc = readback(head_slots, k)) return rc;
    if (head_slots == 3) {
      jy_h.assign(hj(2), hj(2) + m);
      jy_ok = true;
    }
    return MPRGP_OK;
  }
  // Second half, after head_wait: the QPs of BETA (active rows, h = J beta)
  // and of the projection of t2 on the host, both sets of multipliers
  // uploaded with one copy, beta = beta0 - J^T l_b, t2 = v - P(t2),
  // gp = phi + beta, and the sums gp.gp, beta.beta, t2.phi queued into dot
  // results 0-2.  A failure of the projection's QP is kept in spec_fail:
  // it counts only if proportional is evaluated.
  bool spec_fail = false;
  int head_finish(const double *v) {
    if (m > 0) {
      std::vector<double> lb(m, 0.0), lp;
      if (!act.empty()) {
        std::vector<double> ha(act.size());
        for (size_t a = 0; a < act.size(); ++a) ha[a] = hj(0)[act[a]];
        const double tol = qp_tol(ha.data(), (int)ha.size());
        for (const auto &a : ablk) qp_rows(a, hj(0), tol, lb);
      }
      const bool failed = qp_fail;
      proj_lam(hj(1), lp);
      spec_fail = qp_fail && !failed;
      qp_fail = failed;
      stage(0, lb);
      stage(1, lp);
      if (int rc = upload(2)) return rc;
      if (!act.empty())
        if (int rc = jt_apply(0, beta, -1.0, beta)) return rc;
      if (int rc = jt_apply(1, t2, 1.0, t2)) return rc;
    }
    MPRGP_HIP(launch_axpby(s, n, 1.0, phi, 1.0, beta, gp));
    MPRGP_HIP(launch_axpby(s, n, 1.0, v, -1.0, t2, t2));
    return dgroup({{gp, gp}, {beta, beta}, {t2, phi}}, 0);
  }
  // residual and proportional from dot results 0-2 (gamma = 1)
  void head_result(double ab, double *residual, int *prop) const {
    *residual = std::sqrt(hd()[0]);
    const double v2 = hd()[2] * (1.0 / ab);
    const double pp = v2 > 0.0 ? v2 : 0.0;
    *prop = hd()[1] <= pp;
  }
  int precond(bool on) {  // DiagonalInFacePreconSolver::solve (MPRGPPrecondition.h:53-64)
    if (on) MPRGP_HIP(launch_vmul(s, n, phi, dinv, z));
    else MPRGP_HIP(hipMemcpyAsync(z, phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return MPRGP_OK;
  }
};

}  // namespace

namespace mprgp {

int rows_to_internal(const mprgp_matrix *A, int m, const int *j_ptr, const int *j_col, const double *j_val,
                     const double *c, RowsInternal &out) {
  const int n = 3 * A->n_nodes;
  const std::vector<int> &iperm = A->iperm;
  const int nnz = m > 0 ? j_ptr[m] : 0;
  if (m > 0 && (j_ptr[0] != 0 || nnz < 0)) return fail(MPRGP_EBADINPUT, "J row pointers must start at 0");
  out.jc.assign(nnz, 0);
  std::vector<int> tcount(n + 1, 0);
  for (int i = 0; i < m; ++i) {
    if (j_ptr[i + 1] < j_ptr[i]) return fail(MPRGP_EBADINPUT, "J row pointers not monotone at %d", i);
    if (j_ptr[i + 1] == j_ptr[i]) return fail(MPRGP_EBADINPUT, "J row %d is empty", i);
    if (!std::isfinite(c[i])) return fail(MPRGP_EBADINPUT, "c[%d] is not finite", i);
    for (int k = j_ptr[i]; k < j_ptr[i + 1]; ++k) {
      const int col = j_col[k];
      if (col < 0 || col >= n) return fail(MPRGP_EBADINPUT, "J column %d out of range", col);
      if (!std::isfinite(j_val[k])) return fail(MPRGP_EBADINPUT, "J value of row %d is not finite", i);
      out.jc[k] = 3 * iperm[col / 3] + col % 3;
      tcount[out.jc[k] + 1]++;
    }
  }
  for (int i = 0; i < n; ++i) tcount[i + 1] += tcount[i];
  out.tptr = tcount;
  out.tc.assign(nnz, 0);
  out.tv.assign(nnz, 0.0);
  std::vector<int> fill(tcount.begin(), tcount.end() - 1);
  for (int i = 0; i < m; ++i)
    for (int k = j_ptr[i]; k < j_ptr[i + 1]; ++k) {
      const int at = fill[out.jc[k]]++;
      out.tc[at] = i;
      out.tv[at] = j_val[k];
    }
  return MPRGP_OK;
}

}  // namespace mprgp

extern "C" int mprgp_general_solve(mprgp_qp *q, int m, const int *j_ptr, const int *j_col, const double *j_val,
                                   const double *c, const mprgp_params *params, mprgp_stats *stats) {
  MPRGP_TRY
  if (!q || m < 0 || (m > 0 && (!j_ptr || !j_col || !j_val || !c)))
    return fail(MPRGP_EBADINPUT, "bad coupled-row arguments");
  if (m > kMaxGeneralRows) return fail(MPRGP_EBADINPUT, "%d rows; the coupled solve takes at most %d", m, kMaxGeneralRows);
  mprgp_params def;
  if (!params) {
    mprgp_params_default(&def);
    params = &def;
  }
  const mprgp_params &p = *params;
  if (!q->has_rhs || !q->has_x) return fail(MPRGP_ESTATE, "rhs and x must be set");
  if (p.precision != MPRGP_FP64 || p.variant != MPRGP_MONOTONIC)
    return fail(MPRGP_EBADINPUT, "the coupled solve runs the FP64 monotone variant");
  if (!(p.tol == p.tol) || p.max_it < 0) return fail(MPRGP_EBADINPUT, "bad tol/max_it");
  MPRGP_NOT_BUSY(q);
  mprgp_ctx *ctx = q->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  General G;
  G.q = q;
  G.ctx = ctx;
  G.s = ctx->stream;
  G.n = 3 * q->n;
  G.m = m;
  const int n = G.n;
  // J in the internal node order; J^T as a CSR of n rows
  RowsInternal R;
  if (int rc = rows_to_internal(q->A, m, j_ptr, j_col, j_val, c, R)) return rc;
  const int nnz = m > 0 ? j_ptr[m] : 0;
  const std::vector<int> &jc = R.jc, &tptr = R.tptr, &tc = R.tc;
  const std::vector<double> &tv = R.tv;
  G.c.assign(c, c + m);
  if (int rc = G.build_jjt(j_ptr, j_val, jc, tptr, tc)) return rc;
  G.face.assign(m, 0);
  // device block: J, J^T, J v, lambda, dot partials and results
  const size_t parts = ((size_t)n + 255) / 256 + 1;
  const size_t bytes = sizeof(int) * ((size_t)m + 1 + (size_t)nnz + (size_t)n + 1 + (size_t)nnz) +
                       sizeof(double) * (2 * (size_t)nnz + 5 * ((size_t)m + 1) + kMaxDots * parts + 16) + 1024;
  if (int rc = ctx->alloc(&G.blk, bytes)) return rc;
  G.blk_bytes = bytes;
  if (hipHostMalloc((void **)&G.h_lam, sizeof(double) * (5 * ((size_t)m + 1) + 16), hipHostMallocDefault) !=
      hipSuccess) {
    G.h_lam = nullptr;
    return fail(MPRGP_ENOMEM, "pinned staging of %d multipliers", m);
  }
  G.h_rb = G.h_lam + 2 * ((size_t)m + 1);
  {
    char *b8 = (char *)G.blk;
    auto carve = [&](size_t sz) {
      uintptr_t u = ((uintptr_t)b8 + 7) & ~(uintptr_t)7;
      b8 = (char *)(u + sz);
      return (void *)u;
    };
    G.jv = (double *)carve(sizeof(double) * nnz);
    G.tv = (double *)carve(sizeof(double) * nnz);
    G.rb = (double *)carve(sizeof(double) * (3 * ((size_t)m + 1) + 16));
    G.res = G.rb + 3 * ((size_t)m + 1);
    G.lam = (double *)carve(sizeof(double) * 2 * ((size_t)m + 1));
    G.part = (double *)carve(sizeof(double) * kMaxDots * parts);
    G.jp = (int *)carve(sizeof(int) * (m + 1));
    G.jc = (int *)carve(sizeof(int) * nnz);
    G.tp = (int *)carve(sizeof(int) * (n + 1));
    G.tc = (int *)carve(sizeof(int) * nnz);
  }
  hipStream_t s = G.s;
  if (m > 0) {
    MPRGP_HIP(hipMemcpyAsync(G.jp, j_ptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(G.jc, jc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(G.jv, j_val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(G.tc, tc.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(G.tv, tv.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, s));
  }
  MPRGP_HIP(hipMemcpyAsync(G.tp, tptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipStreamSynchronize(s));  // the host arrays above are locals
  G.x = q->x; G.g = q->g; G.p = q->p; G.z = q->z; G.phi = q->phi; G.beta = q->beta; G.y = q->y;
  G.gp = q->ap; G.t = q->py; G.t2 = q->p2; G.t3 = q->beta2; G.tl = q->pu; G.pv = q->pv; G.pw = q->pw;
  G.b = q->b;
  G.dinv = q->A->d_dinv;
  const bool pre = p.preconditioned != 0;
  const double tol = p.tol < 1e-30 ? 1e-30 : p.tol;
  hipEvent_t t0 = ctx->t0, t1 = ctx->t1;
  MPRGP_HIP(hipEventRecord(t0, s));
  // the J/c overload projects x first (MPRGPSolver.h:620-622)
  if (p.project_first) {
    if (int rc = G.project(G.x, G.x)) return rc;
  }
  // alpha_bar: given, or 2/specRad(A, ev, tol) (MPRGPSolver.h:21-32, 42-79)
  double alpha_bar, lmax;
  int piters = 0;
  if (p.alpha_bar > 0.0) {
    alpha_bar = p.alpha_bar;
    lmax = 2.0 / alpha_bar;
  } else {
    const double eps = p.power_tol > 0.0 ? p.power_tol : tol;
    double *tmp = G.pv, *out = G.pw;
    if (p.use_ev && q->ev_valid) {
      MPRGP_HIP(hipMemcpyAsync(tmp, q->ev, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    } else {
      MPRGP_HIP(launch_power_start(s, q->stage, (size_t)n, power_start_norm((size_t)n)));
      MPRGP_HIP(launch_gather_nodes(s, q->stage, tmp, q->A->d_perm, q->n));
    }
    double nout = 1.0;
    int it = 0;
    for (; it <= 1000; ++it) {
      if (int rc = G.spmv(tmp, out)) return rc;
      double d;
      if (int rc = G.dots({{out, out}}, &d)) return rc;
      nout = std::sqrt(d);
      if (nout < kEps) {
        MPRGP_HIP(hipMemcpyAsync(q->ev, tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        q->ev_valid = 1;
        nout = kEps;
        break;
      }
      MPRGP_HIP(launch_axpby(s, n, 1.0 / nout, out, 0.0, out, out));
      // |out/nout - tmp| as the reference forms it: out /= nout, then the difference
      MPRGP_HIP(launch_axpby(s, n, 1.0, out, -1.0, tmp, G.t));
      if (int rc = G.dots({{G.t, G.t}}, &d)) return rc;
      if (std::sqrt(d) <= eps) {
        MPRGP_HIP(hipMemcpyAsync(q->ev, tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        q->ev_valid = 1;
        break;
      }
      std::swap(tmp, out);
    }
    piters = it;
    lmax = nout;
    alpha_bar = 2.0 / nout;
  }
  // MPRGPMonotonic::solve (:412-454).  Every kernel takes the inputs the
  // reference's statement order gives it; decisions that need device values
  // read them back in groups (J v of several vectors, several dot products),
  // and a value that cannot change between two uses is read once: phi at
  // the loop head is the one PHI formed at the end of the previous
  // iteration (same g, same face), J x and J y are kept while x and y are
  // unchanged, and the loop head also forms the CG entry's A p and sums,
  // which no other branch reads.
  double residual = 0.0;
  int iteration = 0, num_cg = 0, num_exp = 0, num_prop = 0;
  // initialize (:148-162)
  if (int rc = G.grad(G.x)) return rc;
  if (int rc = G.face_phi(G.x)) return rc;
  if (int rc = G.precond(pre)) return rc;
  MPRGP_HIP(hipMemcpyAsync(G.p, G.z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  double *AP = G.gp;  // A p shares gp's buffer: gp is read only by its own sum
  for (iteration = 0; iteration < p.max_it;) {
    // computeGradients (:163-173) with phi current, then proportional at x
    if (int rc = G.head_launch(G.x, alpha_bar, m > 0 && G.y_size > 0)) return rc;
    if (int rc = G.head_wait(0)) return rc;
    if (int rc = G.head_finish(G.x)) return rc;
    // the CG entry's A p and its sums p'Ap, z'g (:274-276), formed after gp's sum
    if (int rc = G.spmv(G.p, AP)) return rc;
    if (int rc = G.dgroup({{G.p, AP}, {G.z, G.g}}, 3)) return rc;
    if (int rc = G.readback(0, 5)) return rc;
    int prop = 0;
    G.head_result(alpha_bar, &residual, &prop);
    if (residual <= tol) break;
    if (G.spec_fail) G.qp_fail = true;
    const bool jy_top = G.jy_ok;  // J y of the loop head: y unchanged unless CG steps run
    if (prop) {  // CGMonotonicStep (:272-316)
      G.jx_ok = G.jy_ok = false;
      double alpha = G.hd()[4] / G.hd()[3];
      MPRGP_HIP(launch_axpby(s, n, 1.0, G.x, -alpha, G.p, G.y));
      G.y_size = n;
      if (int rc = G.project(G.y, G.t)) return rc;
      if (int rc = G.func_launch(G.t, 0)) return rc;
      if (int rc = G.func_launch(G.x, 2)) return rc;
      if (int rc = G.readback(0, 4)) return rc;
      double fy = 0.5 * G.hd()[0] - G.hd()[1], fx = 0.5 * G.hd()[2] - G.hd()[3];
      prop = 1;
      while (prop && residual > tol && fy <= fx && iteration < p.max_it) {
        MPRGP_HIP(hipMemcpyAsync(G.x, G.y, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        MPRGP_HIP(launch_axpby(s, n, 1.0, G.g, -alpha, AP, G.g));
        if (int rc = G.compute_phi()) return rc;
        if (int rc = G.precond(pre)) return rc;
        // z'AP and p'AP (beta) come back with the loop head's J products
        if (int rc = G.head_launch(G.x, alpha_bar, false)) return rc;
        if (int rc = G.dgroup({{G.z, AP}, {G.p, AP}}, 3)) return rc;
        if (int rc = G.head_wait(5)) return rc;
        const double beta = G.hd()[3] / G.hd()[4];
        MPRGP_HIP(launch_axpby(s, n, 1.0, G.z, -beta, G.p, G.p));
        if (int rc = G.head_finish(G.x)) return rc;  // gp's sum read before A p overwrites it
        if (int rc = G.spmv(G.p, AP)) return rc;
        if (int rc = G.dgroup({{G.p, AP}, {G.z, G.g}}, 3)) return rc;
        if (int rc = G.readback(0, 5)) return rc;
        G.head_result(alpha_bar, &residual, &prop);
        if (G.spec_fail) G.qp_fail = true;
        alpha = G.hd()[4] / G.hd()[3];
        MPRGP_HIP(launch_axpby(s, n, 1.0, G.x, -alpha, G.p, G.y));
        fx = fy;
        if (int rc = G.project(G.y, G.t)) return rc;
        if (int rc = G.func_launch(G.t, 0)) return rc;
        if (int rc = G.readback(0, 2)) return rc;
        fy = 0.5 * G.hd()[0] - G.hd()[1];
        num_cg++;
        iteration++;
      }
    }
    if (residual <= tol) break;
    bool feas = false;
    if (!prop && G.y_size > 0) {  // isFeasible(y) (:431)
      if (m == 0) {
        feas = true;
      } else if (jy_top && G.jy_ok) {
        feas = G.feasible_h(G.jy_h.data());
      } else {
        if (int rc = G.jmul(G.y)) return rc;
        feas = G.feasible_h(G.hj(0));
      }
    }
    if (prop || G.y_size <= 0 || !feas) {  // ExpMonotonicStep (:250-271)
      num_exp++;
      if (int rc = G.project(G.x, G.y, G.jx_ok ? G.jx_h.data() : nullptr)) return rc;
      G.y_size = n;
      if (int rc = G.grad(G.y)) return rc;
      if (int rc = G.face_phi(G.y)) return rc;
      MPRGP_HIP(launch_axpby(s, n, 1.0, G.y, -alpha_bar, G.phi, G.y));
      if (int rc = G.project(G.y, G.x)) return rc;
      if (int rc = G.grad(G.x)) return rc;
    } else {  // the caller's g = A x - b, then PropStep (:219-243, 435-437)
      num_prop++;
      if (int rc = G.grad(G.x)) return rc;
      double *AD = G.gp;
      if (int rc = G.spmv(G.beta, AD)) return rc;
      double d[2];
      if (int rc = G.dots({{G.beta, AD}, {G.g, G.beta}}, d)) return rc;
      const double alpha = d[1] / d[0];
      MPRGP_HIP(launch_axpby(s, n, 1.0, G.x, -alpha, G.beta, G.t));
      if (int rc = G.project(G.t, G.x)) return rc;
      MPRGP_HIP(launch_axpby(s, n, 1.0, G.g, -alpha, AD, G.g));
    }
    G.jx_ok = G.jy_ok = false;
    if (int rc = G.face_phi(G.x)) return rc;
    if (int rc = G.precond(pre)) return rc;
    MPRGP_HIP(hipMemcpyAsync(G.p, G.z, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    iteration++;
  }
  // the final projection (:442-443)
  if (int rc = G.project(G.x, G.t, G.jx_ok ? G.jx_h.data() : nullptr)) return rc;
  MPRGP_HIP(hipMemcpyAsync(G.x, G.t, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  MPRGP_HIP(hipEventRecord(t1, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  MPRGP_HIP(hipEventElapsedTime(&ms, t0, t1));
  if (G.qp_fail) return fail(MPRGP_EBADINPUT, "a projection QP or the active rows' J J^T factorisation failed "
                                              "(J J^T singular on the rows concerned)");
  const int code = residual <= tol ? 0 : -1;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->code = code;
    stats->iterations = iteration;
    stats->num_cg = num_cg;
    stats->num_exp = num_exp;
    stats->num_prop = num_prop;
    stats->active = (int)G.act.size();
    stats->power_iters = piters;
    stats->ev_valid = q->ev_valid;
    stats->residual = residual;
    stats->alpha_bar = alpha_bar;
    stats->lambda_max = lmax;
    stats->elapsed_ms = ms;
    const double lo = q->A->lambda_min_lower;
    stats->error_bound = -1.0;  // the bound of mprgp_stats holds for decoupled rows
    (void)lo;
  }
  if (!std::isfinite(residual)) return MPRGP_ENAN;
  return code == 0 ? MPRGP_OK : MPRGP_NOT_CONVERGED;
  MPRGP_CATCH
}
