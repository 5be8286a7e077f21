/*
 * mprgp_amd.hpp -- header-only C++ facade with the reference's operator API
 * (src/MPRGPSolver/MPRGPSolver.h:459-500, 536-590 and
 * src/MPRGPSolver/MPRGPProjection.h:320-397) on top of the C ABI in
 * mprgp_amd.h, so that code written against the reference compiles after
 * swapping the include and the namespace:
 *
 *   mprgp_amd::FixedSparseMatrix<double> A(LHS_mat);
 *   mprgp_amd::DecoupledConProjector<double> projector(J, JJt, c);
 *   projector.project(x1, new_pos);
 *   mprgp_amd::MPRGPDecoupledCon<double>::solve<mprgp_amd::FixedSparseMatrix<double>, true>(
 *       A, RHS, projector, new_pos, tol, max_it, "decoupled MPRGP", &power_ev);
 *   mprgp_amd::ICASolver ica(max_it, tol);          // ICAFemSolver's comparator
 *   ica.reset(LHS_mat, RHS);  ica.solve(J, c, new_pos);
 *
 * Types are duck-typed on Eigen's interfaces: a sparse matrix is anything
 * with rows(), cols(), outerIndexPtr(), innerIndexPtr(), valuePtr() and
 * isCompressed() in column-major (Eigen's default) storage; a vector is
 * anything with size(), data() and resize(n) (Eigen::VectorXd,
 * std::vector<double>).  A is converted to row-major storage (a transpose of
 * its index arrays, O(nnz)), so the device multiplies by A itself, bit for
 * bit, even where an assembled A is symmetric only up to rounding.  Only
 * T = double is provided (the simulator's instantiation).
 *
 * Results: the reference's codes (0 converged, -1 not), or a negative
 * MPRGP_E* code for bad input or device errors (mprgp_last_error() says
 * which); the reference asserts instead.  All work runs on device 0 through a
 * context created on first use and kept for the life of the process.
 */
#ifndef MPRGP_AMD_HPP
#define MPRGP_AMD_HPP

#include <cmath>
#include <cstddef>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "mprgp_amd.h"

namespace mprgp_amd {

namespace detail {

inline mprgp_ctx *context() {
  static mprgp_ctx *ctx = [] {
    mprgp_ctx *c = nullptr;
    return mprgp_ctx_create(0, 0, &c) == MPRGP_OK ? c : nullptr;
  }();
  return ctx;
}

// Index arrays of a sparse matrix, copied.  of(): the column-major arrays
// as they are (ptr per column, idx = rows).  rowmajor_of(): the same matrix
// in CSR (ptr per row, idx = columns ascending).
struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> ptr, idx;
  std::vector<double> val;
  template <typename SP>
  static Csr of(const SP &m) {
    Csr c;
    c.rows = (int)m.rows();
    c.cols = (int)m.cols();
    c.ptr.assign(m.outerIndexPtr(), m.outerIndexPtr() + c.cols + 1);
    const int nnz = c.ptr.empty() ? 0 : c.ptr.back();
    c.idx.assign(m.innerIndexPtr(), m.innerIndexPtr() + nnz);
    c.val.assign(m.valuePtr(), m.valuePtr() + nnz);
    return c;
  }
  template <typename SP>
  static Csr rowmajor_of(const SP &m) {
    Csr c;
    c.rows = (int)m.rows();
    c.cols = (int)m.cols();
    const int *op = m.outerIndexPtr();
    const int *ip = m.innerIndexPtr();
    const double *vp = m.valuePtr();
    const int nnz = op[c.cols];
    c.ptr.assign(c.rows + 1, 0);
    for (int k = 0; k < nnz; ++k) c.ptr[ip[k] + 1]++;
    for (int i = 0; i < c.rows; ++i) c.ptr[i + 1] += c.ptr[i];
    c.idx.resize(nnz);
    c.val.resize(nnz);
    std::vector<int> fill(c.ptr.begin(), c.ptr.end() - 1);
    for (int j = 0; j < c.cols; ++j)  // columns in order: each row's columns come out ascending
      for (int k = op[j]; k < op[j + 1]; ++k) {
        const int at = fill[ip[k]]++;
        c.idx[at] = j;
        c.val[at] = vp[k];
      }
    return c;
  }
};

}  // namespace detail

// FixedSparseMatrix<T> (MPRGPSolver.h:459-500): a copy of A, resident on the
// device once first used.
template <typename T = double>
class FixedSparseMatrix;

template <>
class FixedSparseMatrix<double> {
 public:
  template <typename SP>
  explicit FixedSparseMatrix(const SP &M)
      : compressed_(M.isCompressed()) {
    if (compressed_) a_ = detail::Csr::rowmajor_of(M);
  }
  FixedSparseMatrix(const FixedSparseMatrix &o) : a_(o.a_), compressed_(o.compressed_) {}
  ~FixedSparseMatrix() {
    if (dev_) mprgp_matrix_destroy(dev_);
  }
  int rows() const { return a_.rows; }
  // result = A x
  template <typename VEC, typename VEC_OUT>
  int multiply(const VEC &x, VEC_OUT &result) const {
    if (int rc = upload()) return rc;
    result.resize(a_.rows);
    return mprgp_matrix_spmv(dev_, x.data(), result.data());
  }
  // 1/2 x'Ax - x'b
  template <typename VEC, typename VEC_SECOND>
  double funcValue(const VEC &x, const VEC_SECOND &b) const {
    double f = NAN;
    if (upload() == MPRGP_OK) mprgp_matrix_func_value(dev_, x.data(), b.data(), &f);
    return f;
  }
  double diag(const int i) const {
    if (diag_.empty() && upload() == MPRGP_OK) {
      diag_.resize(a_.rows);
      if (mprgp_matrix_diagonal(dev_, diag_.data()) != MPRGP_OK) diag_.clear();
    }
    return i >= 0 && i < (int)diag_.size() ? diag_[i] : NAN;
  }
  const detail::Csr &csr() const { return a_; }

 private:
  int upload() const {
    if (dev_) return MPRGP_OK;
    if (!compressed_ || a_.rows != a_.cols) return MPRGP_EBADINPUT;
    mprgp_ctx *ctx = detail::context();
    if (!ctx) return MPRGP_ENODEV;
    return mprgp_matrix_create_csr(ctx, a_.rows, a_.ptr.data(), a_.idx.data(), a_.val.data(), &dev_);
  }
  detail::Csr a_;
  bool compressed_;
  mutable mprgp_matrix *dev_ = nullptr;
  mutable std::vector<double> diag_;
};

// DecoupledConProjector<T> (MPRGPProjection.h:320-397): the rows J x >= c
// with J J' diagonal, one node per row.  project and isFeasible run on the
// device over a QP that holds only the rows.
template <typename T = double>
class DecoupledConProjector;

template <>
class DecoupledConProjector<double> {
 public:
  // JJt is accepted for the reference's signature; the rows carry |n|^2.
  template <typename SP, typename VEC_JJT, typename VEC_C>
  DecoupledConProjector(const SP &J, const VEC_JJT & /*JJt*/, const VEC_C &c)
      : j_(detail::Csr::of(J)), compressed_(J.isCompressed()), c_(c.data(), c.data() + c.size()) {}
  ~DecoupledConProjector() {
    if (qp_) mprgp_qp_destroy(qp_);
    if (eye_) mprgp_matrix_destroy(eye_);
  }
  DecoupledConProjector(const DecoupledConProjector &) = delete;
  DecoupledConProjector &operator=(const DecoupledConProjector &) = delete;

  template <typename VEC, typename VEC_OUT>
  int project(const VEC &x, VEC_OUT &y) {
    if (int rc = ready((int)x.size())) return rc;
    y.resize(x.size());
    return mprgp_qp_project(qp_, x.data(), y.data());
  }
  template <typename VEC>
  bool isFeasible(const VEC &x) {
    int ok = 0;
    return ready((int)x.size()) == MPRGP_OK && mprgp_qp_is_feasible(qp_, x.data(), &ok) == MPRGP_OK && ok;
  }
  const detail::Csr &con() const { return j_; }
  const std::vector<double> &rhs() const { return c_; }
  bool compressed() const { return compressed_; }

 private:
  int ready(int n) {
    if (qp_ && n == 3 * nodes_) return MPRGP_OK;
    if (!compressed_ || n % 3 != 0 || j_.cols != n || (int)c_.size() != j_.rows) return MPRGP_EBADINPUT;
    mprgp_ctx *ctx = detail::context();
    if (!ctx) return MPRGP_ENODEV;
    if (qp_) mprgp_qp_destroy(qp_);
    if (eye_) mprgp_matrix_destroy(eye_);
    qp_ = nullptr;
    eye_ = nullptr;
    nodes_ = n / 3;
    std::vector<int> rowptr(nodes_ + 1), col(nodes_);
    std::vector<double> blocks(9 * (size_t)nodes_, 0.0);
    for (int i = 0; i < nodes_; ++i) {
      rowptr[i + 1] = i + 1;
      col[i] = i;
      blocks[9 * (size_t)i] = blocks[9 * (size_t)i + 4] = blocks[9 * (size_t)i + 8] = 1.0;
    }
    if (int rc = mprgp_matrix_create_bsr3(ctx, nodes_, rowptr.data(), col.data(), blocks.data(), &eye_)) return rc;
    if (int rc = mprgp_qp_create(eye_, &qp_)) return rc;
    return mprgp_qp_set_rows_sparse(qp_, j_.rows, 1, j_.ptr.data(), j_.idx.data(), j_.val.data(), c_.data());
  }
  detail::Csr j_;
  bool compressed_;
  std::vector<double> c_;
  int nodes_ = 0;
  mprgp_matrix *eye_ = nullptr;
  mprgp_qp *qp_ = nullptr;
};

namespace detail {

template <typename VEC>
struct EvIo {  // the reference's Vec* ev: a start vector when its size is n, written back on an early stop
  std::vector<double> buf;
  int valid = 0;
  EvIo(VEC *ev, int n) : buf(n, 0.0) {
    if (ev && (int)ev->size() == n) {
      buf.assign(ev->data(), ev->data() + n);
      valid = 1;
    }
  }
  void back(VEC *ev, int valid_out) {
    if (!ev || !valid_out) return;
    ev->resize(buf.size());
    for (size_t i = 0; i < buf.size(); ++i) ev->data()[i] = buf[i];
  }
};

template <typename MAT>
const Csr &matrix_csr(const MAT &A, Csr &tmp) {
  tmp = Csr::rowmajor_of(A);
  return tmp;
}
inline const Csr &matrix_csr(const FixedSparseMatrix<double> &A, Csr &) { return A.csr(); }

template <typename MAT>
bool matrix_compressed(const MAT &A) {
  return A.isCompressed();
}
inline bool matrix_compressed(const FixedSparseMatrix<double> &) { return true; }

// One solve through the object API: rows from (J, c), x read and written.
template <typename VEC>
int solve_rows(const Csr &a, const double *B, const Csr &j, const double *c, VEC &x, double tol, int max_it,
               int project_first, bool preconditioned, VEC *ev) {
  mprgp_ctx *ctx = context();
  if (!ctx) return MPRGP_ENODEV;
  if (a.rows != a.cols || (int)x.size() != a.rows || j.cols != a.rows) return MPRGP_EBADINPUT;
  mprgp_matrix *m = nullptr;
  mprgp_qp *q = nullptr;
  int rc = mprgp_matrix_create_csr(ctx, a.rows, a.ptr.data(), a.idx.data(), a.val.data(), &m);
  if (!rc) rc = mprgp_qp_create(m, &q);
  if (!rc) rc = mprgp_qp_set_rows_sparse(q, j.rows, 1, j.ptr.data(), j.idx.data(), j.val.data(), c);
  if (!rc) rc = mprgp_qp_set_rhs(q, B);
  if (!rc) rc = mprgp_qp_set_x(q, x.data());
  EvIo<VEC> evio(ev, a.rows);
  if (!rc && evio.valid) rc = mprgp_qp_set_ev(q, evio.buf.data());
  int code = rc;
  if (!rc) {
    mprgp_params p;
    mprgp_params_default(&p);
    p.tol = tol;
    p.max_it = max_it;
    p.project_first = project_first;
    p.preconditioned = preconditioned ? 1 : 0;
    p.use_ev = evio.valid;
    mprgp_stats st;
    code = mprgp_qp_solve(q, &p, &st);
    if (code == MPRGP_OK || code == MPRGP_NOT_CONVERGED) {
      int valid_out = 0;
      if (mprgp_qp_get_x(q, x.data()) != MPRGP_OK) code = MPRGP_EHIP;
      else if (mprgp_qp_get_ev(q, evio.buf.data(), &valid_out) == MPRGP_OK) evio.back(ev, valid_out);
    }
  }
  if (q) mprgp_qp_destroy(q);
  if (m) mprgp_matrix_destroy(m);
  return code;
}

}  // namespace detail

// MPRGPDecoupledCon<T> (MPRGPSolver.h:536-590).
template <typename T = double>
class MPRGPDecoupledCon;

template <>
class MPRGPDecoupledCon<double> {
 public:
  // Projector overload (:543-559): x is not projected first.
  template <typename MAT, bool preconditioned = true, typename VEC, typename VEC_B>
  static int solve(const MAT &A, const VEC_B &B, DecoupledConProjector<double> &projector, VEC &x,
                   const double tol = 1e-3, const int max_it = 1000, const std::string solver_name = "MPRGP",
                   VEC *ev = NULL) {
    (void)solver_name;
    if (!detail::matrix_compressed(A) || !projector.compressed()) return MPRGP_EBADINPUT;
    detail::Csr tmp;
    const detail::Csr &a = detail::matrix_csr(A, tmp);
    if ((int)B.size() != a.rows) return MPRGP_EBADINPUT;
    return detail::solve_rows(a, B.data(), projector.con(), projector.rhs().data(), x, tol, max_it, 0,
                              preconditioned, ev);
  }

  // J/c overload (:561-574): x is projected first.
  template <typename MAT, bool preconditioned = true, typename VEC, typename VEC_B, typename SP, typename VEC_C>
  static int solve(const MAT &A, const VEC_B &B, const SP &J, const VEC_C &c, VEC &x, const double tol = 1e-3,
                   const int max_it = 1000, const std::string solver_name = "MPRGP", VEC *ev = NULL) {
    (void)solver_name;
    if (!detail::matrix_compressed(A) || !J.isCompressed()) return MPRGP_EBADINPUT;
    detail::Csr tmp;
    const detail::Csr &a = detail::matrix_csr(A, tmp);
    const detail::Csr j = detail::Csr::of(J);
    if ((int)B.size() != a.rows || (int)c.size() != j.rows) return MPRGP_EBADINPUT;
    return detail::solve_rows(a, B.data(), j, c.data(), x, tol, max_it, 1, preconditioned, ev);
  }

  // File overload (:576-589): the QP written by writeQP; x receives its
  // solution (resized to the QP's size).
  template <bool preconditioned = true, typename VEC>
  static int solve(const std::string file_name, VEC &x, const double tol = 1e-3, const int max_it = 1000,
                   const std::string solver_name = "MPRGP", VEC *ev = NULL) {
    (void)solver_name;
    static_assert(preconditioned, "the file overload runs the preconditioned solve");
    int n = 0, m = 0;
    long long annz = 0, jnnz = 0;
    if (int rc = mprgp_qp_file_info(file_name.c_str(), &n, &m, &annz, &jnnz)) return rc;
    x.resize(n);
    detail::EvIo<VEC> evio(ev, n);
    int ev_valid = evio.valid;
    int n_out = 0;
    mprgp_stats st;
    const int code = mprgp_decoupled_solve_file(0, file_name.c_str(), x.data(), n, &n_out, tol, max_it,
                                                evio.buf.data(), &ev_valid, &st);
    if (code == MPRGP_OK || code == MPRGP_NOT_CONVERGED) evio.back(ev, ev_valid);
    return code;
  }
};

// ICASolver (ICASolver.h:287-463): the comparator, on the device through
// mprgp_ica_solve.  reset(A, r) keeps a copy of A and r; solve(J, p, x) runs
// from x and writes the result into it.
class ICASolver {
 public:
  ICASolver(const int max_it = 1000, const double tol = 1e-4, const int gs_max_it = 20,
            const double gs_tol = 1e-3)
      : max_it_(max_it), tol_(tol), gs_max_it_(gs_max_it), gs_tol_(gs_tol) {}

  template <typename SP, typename VEC>
  void reset(const SP &A, const VEC &r) {
    a_ = detail::Csr::rowmajor_of(A);
    r_.assign(r.data(), r.data() + r.size());
  }
  template <typename SP>
  void reset(const SP &A) {
    a_ = detail::Csr::rowmajor_of(A);
    r_.assign(a_.rows, 0.0);
  }

  // true when |A x - r - J^T lambda| <= tol was reached; a negative MPRGP_E*
  // code leaves iterations at -1 and returns false
  template <typename SP, typename VEC_P, typename VEC>
  bool solve(const SP &J, const VEC_P &p, VEC &x) {
    iterations_ = -1;
    residual_ = -1.0;
    status_ = MPRGP_EBADINPUT;
    const detail::Csr j = detail::Csr::rowmajor_of(J);
    if (a_.rows == 0 || a_.rows != a_.cols || (int)x.size() != a_.rows || (int)r_.size() != a_.rows ||
        j.cols != a_.rows || (int)p.size() != j.rows)
      return false;
    mprgp_ctx *ctx = detail::context();
    if (!ctx) {
      status_ = MPRGP_ENODEV;
      return false;
    }
    mprgp_matrix *m = nullptr;
    mprgp_qp *q = nullptr;
    int rc = mprgp_matrix_create_csr(ctx, a_.rows, a_.ptr.data(), a_.idx.data(), a_.val.data(), &m);
    if (!rc) rc = mprgp_qp_create(m, &q);
    if (!rc) rc = mprgp_qp_set_rhs(q, r_.data());
    if (!rc) rc = mprgp_qp_set_x(q, x.data());
    lambda_.assign(j.rows, 0.0);
    if (!rc) {
      mprgp_stats st;
      int sweeps = 0;
      rc = mprgp_ica_solve(q, j.rows, j.ptr.data(), j.idx.data(), j.val.data(), p.data(), tol_, max_it_,
                           gs_max_it_, gs_tol_, lambda_.data(), &sweeps, &st);
      if (rc == MPRGP_OK || rc == MPRGP_NOT_CONVERGED) {
        iterations_ = st.iterations;
        residual_ = st.residual;
        if (mprgp_qp_get_x(q, x.data()) != MPRGP_OK) rc = MPRGP_EHIP;
      }
    }
    if (q) mprgp_qp_destroy(q);
    if (m) mprgp_matrix_destroy(m);
    status_ = rc;
    return rc == MPRGP_OK;
  }

  double getResidual() const { return residual_; }
  int getIterations() const { return iterations_; }
  const std::vector<double> &getLambda() const { return lambda_; }
  int status() const { return status_; }  // the last solve's mprgp status code

 private:
  int max_it_;
  double tol_;
  int gs_max_it_;
  double gs_tol_;
  detail::Csr a_;
  std::vector<double> r_, lambda_;
  int iterations_ = -1;
  double residual_ = -1.0;
  int status_ = MPRGP_OK;
};

}  // namespace mprgp_amd

#endif  // MPRGP_AMD_HPP
