// mprgp_qpfile.cpp -- the reference's binary QP file (A x = B subject to
// J x >= c, initial value x0), read into CSR arrays and solved from file.
//
// Layout (writeQP / loadQP, MPRGPUtility.h:406-487), host byte order:
//   A:  size_t rows, size_t cols, size_t nnz, then nnz triplets
//       {int32 row, int32 col, double value} (Eigen::Triplet<double>, 16 B)
//   B:  rows doubles
//   J:  size_t rows, size_t cols, size_t nnz, then nnz triplets
//   c:  J.rows doubles
//   x0: A.rows doubles
// Duplicate triplets add up, as Eigen's setFromTriplets does.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/mprgp_amd.h"
#include "mprgp_error.h"
#include "mprgp_qpfile.h"

namespace {

struct Triplet {  // Eigen::Triplet<double, int>
  int32_t row, col;
  double value;
};
static_assert(sizeof(Triplet) == 16, "Eigen::Triplet<double> is 16 bytes");

struct Sparse {
  size_t rows = 0, cols = 0;
  std::vector<Triplet> t;
};

struct QpFile {
  Sparse A, J;
  std::vector<double> B, c, x0;
};

struct FileCloser {
  void operator()(FILE *f) const { if (f) std::fclose(f); }
};

// Bytes from the current position to the end of the file (or 0 when the
// stream cannot seek): every count a header states is checked against it
// before anything is sized by that count.
size_t bytes_left(FILE *f) {
  const long pos = std::ftell(f);
  if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return 0;
  const long end = std::ftell(f);
  if (end < pos || std::fseek(f, pos, SEEK_SET) != 0) return 0;
  return (size_t)(end - pos);
}

bool read_sparse(FILE *f, Sparse &s) {
  size_t hdr[3];
  if (std::fread(hdr, sizeof(size_t), 3, f) != 3) return false;
  s.rows = hdr[0];
  s.cols = hdr[1];
  const size_t nnz = hdr[2];
  if (s.rows > 0x7fffffff || s.cols > 0x7fffffff || nnz > bytes_left(f) / sizeof(Triplet)) return false;
  s.t.resize(nnz);
  if (nnz && std::fread(s.t.data(), sizeof(Triplet), nnz, f) != nnz) return false;
  for (const Triplet &e : s.t)
    if (e.row < 0 || e.col < 0 || (size_t)e.row >= s.rows || (size_t)e.col >= s.cols) return false;
  return true;
}

bool read_doubles(FILE *f, std::vector<double> &v, size_t n) {
  if (n > bytes_left(f) / sizeof(double)) return false;
  v.resize(n);
  return n == 0 || std::fread(v.data(), sizeof(double), n, f) == n;
}

int load(const char *file_name, QpFile &q) {
  if (!file_name) return mprgp::fail(MPRGP_EBADINPUT, "NULL file name");
  std::unique_ptr<FILE, FileCloser> f(std::fopen(file_name, "rb"));
  if (!f) return mprgp::fail(MPRGP_EBADINPUT, "cannot open QP file %s", file_name);
  if (!read_sparse(f.get(), q.A)) return mprgp::fail(MPRGP_EBADINPUT, "%s: bad matrix A", file_name);
  if (q.A.rows != q.A.cols) return mprgp::fail(MPRGP_EBADINPUT, "%s: A is not square", file_name);
  if (!read_doubles(f.get(), q.B, q.A.rows)) return mprgp::fail(MPRGP_EBADINPUT, "%s: short B", file_name);
  if (!read_sparse(f.get(), q.J)) return mprgp::fail(MPRGP_EBADINPUT, "%s: bad matrix J", file_name);
  if (q.J.cols != q.A.rows)
    return mprgp::fail(MPRGP_EBADINPUT, "%s: J has %zu columns, A %zu rows", file_name, q.J.cols, q.A.rows);
  if (!read_doubles(f.get(), q.c, q.J.rows)) return mprgp::fail(MPRGP_EBADINPUT, "%s: short c", file_name);
  if (!read_doubles(f.get(), q.x0, q.A.rows)) return mprgp::fail(MPRGP_EBADINPUT, "%s: short x0", file_name);
  return MPRGP_OK;
}

// CSR of a triplet list: rows in order, columns ascending, duplicates summed.
void to_csr(const Sparse &s, std::vector<int> &ptr, std::vector<int> &col, std::vector<double> &val) {
  std::vector<Triplet> t = s.t;
  std::stable_sort(t.begin(), t.end(), [](const Triplet &a, const Triplet &b) {
    return a.row != b.row ? a.row < b.row : a.col < b.col;
  });
  ptr.assign(s.rows + 1, 0);
  col.clear();
  val.clear();
  for (size_t k = 0; k < t.size(); ++k) {
    if (!col.empty() && k > 0 && t[k].row == t[k - 1].row && t[k].col == t[k - 1].col) {
      val.back() += t[k].value;
      continue;
    }
    col.push_back(t[k].col);
    val.push_back(t[k].value);
    ptr[t[k].row + 1]++;
  }
  for (size_t r = 0; r < s.rows; ++r) ptr[r + 1] += ptr[r];
}

bool write_sparse(FILE *f, size_t rows, size_t cols, const int *ptr, const int *col, const double *val) {
  const size_t nnz = rows ? (size_t)ptr[rows] : 0;
  const size_t hdr[3] = {rows, cols, nnz};
  if (std::fwrite(hdr, sizeof(size_t), 3, f) != 3) return false;
  std::vector<Triplet> t;
  t.reserve(nnz);
  for (size_t r = 0; r < rows; ++r)
    for (int k = ptr[r]; k < ptr[r + 1]; ++k) t.push_back(Triplet{(int32_t)r, col[k], val[k]});
  return nnz == 0 || std::fwrite(t.data(), sizeof(Triplet), nnz, f) == nnz;
}

bool write_doubles(FILE *f, const double *v, size_t n) {
  return n == 0 || std::fwrite(v, sizeof(double), n, f) == n;
}

}  // namespace

extern "C" int mprgp_qp_file_info(const char *file_name, int *n, int *m, long long *a_nnz, long long *j_nnz) {
  MPRGP_TRY
  if (!n || !m || !a_nnz || !j_nnz) return mprgp::fail(MPRGP_EBADINPUT, "NULL argument");
  QpFile q;
  if (int rc = load(file_name, q)) return rc;
  std::vector<int> ptr, col;
  std::vector<double> val;
  to_csr(q.A, ptr, col, val);
  *a_nnz = (long long)col.size();
  to_csr(q.J, ptr, col, val);
  *j_nnz = (long long)col.size();
  *n = (int)q.A.rows;
  *m = (int)q.J.rows;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_file_read(const char *file_name, int *a_ptr, int *a_col, double *a_val, double *B,
                                  int *j_ptr, int *j_col, double *j_val, double *c, double *x0) {
  MPRGP_TRY
  QpFile q;
  if (int rc = load(file_name, q)) return rc;
  if (!a_ptr || !a_col || !a_val || !B || !j_ptr || !x0 || (q.J.rows && !c) ||
      (!q.J.t.empty() && (!j_col || !j_val)))
    return mprgp::fail(MPRGP_EBADINPUT, "NULL argument");
  std::vector<int> ptr, col;
  std::vector<double> val;
  to_csr(q.A, ptr, col, val);
  std::copy(ptr.begin(), ptr.end(), a_ptr);
  std::copy(col.begin(), col.end(), a_col);
  std::copy(val.begin(), val.end(), a_val);
  to_csr(q.J, ptr, col, val);
  std::copy(ptr.begin(), ptr.end(), j_ptr);
  std::copy(col.begin(), col.end(), j_col);
  std::copy(val.begin(), val.end(), j_val);
  std::copy(q.B.begin(), q.B.end(), B);
  std::copy(q.c.begin(), q.c.end(), c);
  std::copy(q.x0.begin(), q.x0.end(), x0);
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_write_qp(const char *file_name, int n, const int *a_ptr, const int *a_col,
                              const double *a_val, const double *B, int m, const int *j_ptr, const int *j_col,
                              const double *j_val, const double *c, const double *x0) {
  MPRGP_TRY
  if (!file_name || n < 0 || m < 0 || !a_ptr || !B || !x0 || !j_ptr || (m > 0 && !c))
    return mprgp::fail(MPRGP_EBADINPUT, "bad QP file arguments");
  std::unique_ptr<FILE, FileCloser> f(std::fopen(file_name, "wb"));
  if (!f) return mprgp::fail(MPRGP_EBADINPUT, "cannot create QP file %s", file_name);
  const bool ok = write_sparse(f.get(), n, n, a_ptr, a_col, a_val) && write_doubles(f.get(), B, n) &&
                  write_sparse(f.get(), m, n, j_ptr, j_col, j_val) && write_doubles(f.get(), c, m) &&
                  write_doubles(f.get(), x0, n);
  if (!ok) return mprgp::fail(MPRGP_EBADINPUT, "writing QP file %s failed", file_name);
  return MPRGP_OK;
  MPRGP_CATCH
}

// The file as CSR arrays for the solve-from-file overload.
int mprgp::read_qp_csr(const char *file_name, QpCsr &out) {
  QpFile q;
  if (int rc = load(file_name, q)) return rc;
  out.n = (int)q.A.rows;
  out.m = (int)q.J.rows;
  to_csr(q.A, out.ap, out.ac, out.av);
  to_csr(q.J, out.jp, out.jc, out.jv);
  out.B = std::move(q.B);
  out.c = std::move(q.c);
  out.x0 = std::move(q.x0);
  return MPRGP_OK;
}
