// qpfile_san.cpp -- reads every file named on the command line through the
// C ABI's QP-file entry points (mprgp_qp_file_info, then mprgp_qp_file_read
// into arrays of the reported sizes), for a build of the reader under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py).
// Prints "<status> <file>" per file; any memory or UB error aborts.
#include <cstdio>
#include <vector>

#include "../../include/mprgp_amd.h"

int main(int argc, char **argv) {
  for (int i = 1; i < argc; ++i) {
    int n = 0, m = 0;
    long long a_nnz = 0, j_nnz = 0;
    int rc = mprgp_qp_file_info(argv[i], &n, &m, &a_nnz, &j_nnz);
    if (rc == MPRGP_OK) {
      std::vector<int> ap(n + 1), ac(a_nnz), jp(m + 1), jc(j_nnz);
      std::vector<double> av(a_nnz), B(n), jv(j_nnz), c(m), x0(n);
      rc = mprgp_qp_file_read(argv[i], ap.data(), ac.data(), av.data(), B.data(), jp.data(), jc.data(),
                              jv.data(), c.data(), x0.data());
    }
    std::printf("%d %s\n", rc, argv[i]);
  }
  return 0;
}
