// mprgp_qpfile.h -- the reference's binary QP file as CSR arrays (internal
// header; host-only, no HIP dependency).
#pragma once
#include <vector>

namespace mprgp {

// A (n x n) and J (m x n) in CSR with columns ascending and duplicate
// triplets summed, B (n), c (m), x0 (n).
struct QpCsr {
  int n = 0, m = 0;
  std::vector<int> ap, ac, jp, jc;
  std::vector<double> av, jv, B, c, x0;
};

// MPRGP_OK, or MPRGP_EBADINPUT for a file that cannot be read or is malformed.
int read_qp_csr(const char *file_name, QpCsr &out);

}  // namespace mprgp
