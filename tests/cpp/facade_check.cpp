// Exercises include/mprgp_amd.hpp the way the reference's callers use
// MPRGPDecoupledCon, DecoupledConProjector and FixedSparseMatrix, with small
// stand-ins for Eigen's column-major SparseMatrix and VectorXd (Eigen is not
// available here).  Reads a QP file, runs the three solve overloads and the
// projector/matrix methods, and prints one line per result for the test that
// runs it:  name code v0 v1 ... (values %.17g).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "mprgp_amd.hpp"

struct Vec {
  std::vector<double> v;
  size_t size() const { return v.size(); }
  double *data() { return v.data(); }
  const double *data() const { return v.data(); }
  void resize(size_t n) { v.resize(n); }
};

// column-major sparse: outer = columns
struct Sparse {
  int r = 0, c = 0;
  std::vector<int> outer, inner;
  std::vector<double> val;
  int rows() const { return r; }
  int cols() const { return c; }
  const int *outerIndexPtr() const { return outer.data(); }
  const int *innerIndexPtr() const { return inner.data(); }
  const double *valuePtr() const { return val.data(); }
  bool isCompressed() const { return true; }
};

// CSR (rows x cols) -> column-major arrays of the same matrix
static Sparse from_csr(int rows, int cols, const std::vector<int> &ptr, const std::vector<int> &idx,
                       const std::vector<double> &val) {
  Sparse s;
  s.r = rows;
  s.c = cols;
  s.outer.assign(cols + 1, 0);
  for (int k = 0; k < ptr[rows]; ++k) s.outer[idx[k] + 1]++;
  for (int j = 0; j < cols; ++j) s.outer[j + 1] += s.outer[j];
  s.inner.resize(ptr[rows]);
  s.val.resize(ptr[rows]);
  std::vector<int> fill(s.outer.begin(), s.outer.end() - 1);
  for (int i = 0; i < rows; ++i)
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int at = fill[idx[k]]++;
      s.inner[at] = i;
      s.val[at] = val[k];
    }
  return s;
}

static void emit(const char *name, int code, const Vec &x) {
  std::printf("%s %d", name, code);
  for (double v : x.v) std::printf(" %.17g", v);
  std::printf("\n");
}

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s qp_file tol max_it\n", argv[0]);
    return 2;
  }
  const std::string file = argv[1];
  const double tol = std::atof(argv[2]);
  const int max_it = std::atoi(argv[3]);
  int n = 0, m = 0;
  long long annz = 0, jnnz = 0;
  if (mprgp_qp_file_info(file.c_str(), &n, &m, &annz, &jnnz) != MPRGP_OK) {
    std::fprintf(stderr, "read: %s\n", mprgp_last_error());
    return 1;
  }
  std::vector<int> ap(n + 1), ai(annz), jp(m + 1), ji(jnnz);
  std::vector<double> av(annz), jv(jnnz);
  Vec B, c, x0;
  B.resize(n);
  c.resize(m);
  x0.resize(n);
  if (mprgp_qp_file_read(file.c_str(), ap.data(), ai.data(), av.data(), B.data(), jp.data(), ji.data(), jv.data(),
                         c.data(), x0.data()) != MPRGP_OK) {
    std::fprintf(stderr, "read: %s\n", mprgp_last_error());
    return 1;
  }
  const Sparse A = from_csr(n, n, ap, ai, av);
  const Sparse J = from_csr(m, n, jp, ji, jv);
  Vec JJt;  // unused by the facade, kept for the reference's signature
  JJt.resize(m);

  // DecoupledMprgpFemSolver::forward's sequence (MprgpFemSolver.cpp:514-533)
  const mprgp_amd::FixedSparseMatrix<double> FA(A);
  mprgp_amd::DecoupledConProjector<double> projector(J, JJt, c);
  Vec y;
  const int pc = projector.project(x0, y);
  emit("project", pc, y);
  std::printf("feasible %d %d\n", projector.isFeasible(x0) ? 1 : 0, projector.isFeasible(y) ? 1 : 0);
  Vec ev;  // empty: the power method starts from its own vector, as with an empty power_ev
  const int c1 = mprgp_amd::MPRGPDecoupledCon<double>::solve<mprgp_amd::FixedSparseMatrix<double>, true>(
      FA, B, projector, y, tol, max_it, "decoupled MPRGP", &ev);
  emit("projector_overload", c1, y);
  std::printf("ev_size %zu\n", ev.size());

  // the J/c overload from the unprojected start
  Vec x = x0;
  const int c2 = mprgp_amd::MPRGPDecoupledCon<double>::solve<Sparse, true>(A, B, J, c, x, tol, max_it);
  emit("jc_overload", c2, x);

  // the file overload
  Vec xf;
  const int c3 = mprgp_amd::MPRGPDecoupledCon<double>::solve(file, xf, tol, max_it);
  emit("file_overload", c3, xf);

  // the matrix concept at the solution
  Vec ax;
  const int c4 = FA.multiply(x, ax);
  emit("multiply", c4, ax);
  std::printf("funcValue %.17g\n", FA.funcValue(x, B));
  Vec d;
  d.resize(n);
  for (int i = 0; i < n; ++i) d.v[i] = FA.diag(i);
  emit("diag", 0, d);

  // the ICA comparator as ICAFemSolver::forward calls it, from the start x0
  mprgp_amd::ICASolver ica(max_it, tol);
  ica.reset(A, B);
  Vec xi = x0;
  const bool ok = ica.solve(J, c, xi);
  emit("ica", ok ? 0 : -1, xi);
  Vec lam;
  lam.v = ica.getLambda();
  emit("ica_lambda", ica.getIterations(), lam);
  return 0;
}
