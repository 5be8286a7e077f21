// graph_churn.cpp -- the batch driver's captured-graph life cycle under
// host AddressSanitizer (tests/test_gpu_host_asan.py runs it against a
// library built with -Xarch_host -fsanitize=address).
//
// The sequence is the autotuner's: the same batch solved with each candidate
// launch shape in turn (wavefronts per workgroup 4, 2, 2, 2, 4, 8, 1 with
// 16, 16, 32, 8, 32, 16, 16 launches per graph), three times over, so the
// context's tile table regrows (one wavefront per workgroup needs the most
// tiles) and its captured graphs are destroyed and captured again; then a
// lone QP (direct and cached-graph paths) and an engine that admits and
// retires QPs of several sizes.  Every solve must converge (the mixed form's
// refinement stalls on the longer chains, whose rows carry about 7 units of
// force each, and the FP64 form finishes them); any use-after-free or
// overflow in the host code aborts the run.
//
// Matrices: per QP a chain of n nodes, A = (1 + h^2 k) diag blocks with
// -h^2 k/2 couplings to both neighbours (SPD, diagonally dominant); one
// half-space row y >= 0 on every 7th node; b pulls the chain downwards.
// alpha_bar = 2/(1 + 2k), the Gershgorin bound on |A| (the power method's
// estimate from below converges slowly on a long chain's clustered top
// eigenvalues).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <unistd.h>

#include "../../include/mprgp_amd.h"

#define CHECK(call)                                                              \
  do {                                                                           \
    int rc_ = (call);                                                            \
    if (rc_ != MPRGP_OK) {                                                       \
      std::fprintf(stderr, "%s -> %d: %s\n", #call, rc_, mprgp_last_error());     \
      std::exit(1);                                                              \
    }                                                                            \
  } while (0)

constexpr double kChain = 40.0;
constexpr double kAlphaBar = 2.0 / (1.0 + 2.0 * kChain);

static mprgp_qp *chain_qp(mprgp_ctx *ctx, int n, mprgp_matrix **Aout) {
  std::vector<int> rowptr(n + 1), col;
  std::vector<double> blocks;
  const double k = kChain;
  for (int i = 0; i < n; ++i) {
    rowptr[i] = (int)col.size();
    for (int j = i - 1; j <= i + 1; ++j) {
      if (j < 0 || j >= n) continue;
      col.push_back(j);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) blocks.push_back(r != c ? 0.0 : (j == i ? 1.0 + k : -0.5 * k));
    }
  }
  rowptr[n] = (int)col.size();
  mprgp_matrix *A = nullptr;
  CHECK(mprgp_matrix_create_bsr3(ctx, n, rowptr.data(), col.data(), blocks.data(), &A));
  mprgp_qp *q = nullptr;
  CHECK(mprgp_qp_create(A, &q));
  std::vector<double> b(3 * (size_t)n), x(3 * (size_t)n);
  std::vector<int> node;
  std::vector<double> nrm, c;
  for (int i = 0; i < n; ++i) {
    b[3 * i] = 0.01 * std::sin(0.1 * i);
    b[3 * i + 1] = -1.0;
    b[3 * i + 2] = 0.01 * std::cos(0.1 * i);
    x[3 * i + 1] = 0.05;
    if (i % 7 == 0) {
      node.push_back(i);
      nrm.insert(nrm.end(), {0.0, 1.0, 0.0});
      c.push_back(0.0);
    }
  }
  CHECK(mprgp_qp_set_rows(q, (int)node.size(), node.data(), nrm.data(), c.data()));
  CHECK(mprgp_qp_set_rhs(q, b.data()));
  CHECK(mprgp_qp_set_x(q, x.data()));
  *Aout = A;
  return q;
}

int main() {
  mprgp_ctx *ctx = nullptr;
  CHECK(mprgp_ctx_create(0, 0, &ctx));
  const int sizes[] = {900, 2500, 7000, 12000, 20000, 31000, 4100, 16000};
  std::vector<mprgp_qp *> qps;
  std::vector<mprgp_matrix *> mats;
  for (int n : sizes) {
    mprgp_matrix *A = nullptr;
    qps.push_back(chain_qp(ctx, n, &A));
    mats.push_back(A);
  }
  const int shapes[][2] = {{4, 16}, {2, 16}, {2, 32}, {2, 8}, {4, 32}, {8, 16}, {1, 16}};
  std::vector<mprgp_stats> st(qps.size());
  long solves = 0;
  for (int round = 0; round < 3; ++round)
    for (const auto &sh : shapes)
      for (int prec = 0; prec < 2; ++prec) {
        std::vector<mprgp_params> p(qps.size());
        for (size_t i = 0; i < qps.size(); ++i) {
          mprgp_params_default(&p[i]);
          p[i].tol = 1e-6;
          p[i].max_it = 20000;
          p[i].alpha_bar = kAlphaBar;
          p[i].precision = prec ? MPRGP_MIXED : MPRGP_FP64;
          p[i].project_first = 1;
          p[i].tiles = sh[0];
          p[i].graph_steps = sh[1];
          CHECK(mprgp_qp_reset_x(qps[i]));
        }
        // all QPs, then every other one (a smaller table with the same graphs)
        const int rc = mprgp_solve_batch(ctx, (int)qps.size(), qps.data(), p.data(), 1, st.data());
        if (rc != MPRGP_OK) {
          for (size_t i = 0; i < qps.size(); ++i)
            std::fprintf(stderr,
                         "qp %zu (%d nodes) shape %d/%d %s: code %d, %d iterations, %d sweeps, fp64 finish %d, "
                         "residual %g\n",
                         i, sizes[i], sh[0], sh[1], prec ? "mixed" : "fp64", st[i].code, st[i].iterations,
                         st[i].refinements, st[i].fp64_finish, st[i].residual);
          CHECK(rc);
        }
        std::vector<mprgp_qp *> half;
        std::vector<mprgp_params> ph;
        for (size_t i = 0; i < qps.size(); i += 2) {
          half.push_back(qps[i]);
          ph.push_back(p[i]);
          CHECK(mprgp_qp_reset_x(qps[i]));
        }
        CHECK(mprgp_solve_batch(ctx, (int)half.size(), half.data(), ph.data(), 1, st.data()));
        solves += (long)qps.size() + (long)half.size();
        std::printf("round %d shape %d/%d %s: ok\n", round, sh[0], sh[1], prec ? "mixed" : "fp64");
        std::fflush(stdout);
      }
  // a lone QP: direct launches and cached graphs
  for (int g : {-16, 16, 32}) {
    mprgp_params p;
    mprgp_params_default(&p);
    p.tol = 1e-6;
    p.max_it = 20000;
    p.alpha_bar = kAlphaBar;
    p.graph_steps = g;
    CHECK(mprgp_qp_reset_x(qps[3]));
    CHECK(mprgp_qp_solve(qps[3], &p, &st[0]));
    ++solves;
  }
  // engines of two shapes admitting and retiring QPs of every size
  for (int tiles : {4, 1}) {
    mprgp_engine_config cfg;
    mprgp_engine_config_default(&cfg);
    cfg.max_active = 3;
    cfg.tiles = tiles;
    mprgp_engine *eng = nullptr;
    CHECK(mprgp_engine_create(ctx, &cfg, &eng));
    for (int rep = 0; rep < 2; ++rep) {
      std::vector<long long> tickets;
      for (mprgp_qp *q : qps) {
        mprgp_params p;
        mprgp_params_default(&p);
        p.tol = 1e-6;
        p.max_it = 20000;
        p.alpha_bar = kAlphaBar;
        p.precision = rep ? MPRGP_MIXED : MPRGP_FP64;
        CHECK(mprgp_qp_reset_x(q));
        long long t = 0;
        CHECK(mprgp_engine_submit(eng, q, &p, &t));
        tickets.push_back(t);
      }
      for (long long t : tickets) CHECK(mprgp_engine_wait(eng, t, -1, &st[0]));
      solves += (long)tickets.size();
    }
    mprgp_engine_destroy(eng);
    std::printf("engine with %d wavefronts per workgroup: ok\n", tiles);
    std::fflush(stdout);
  }
  for (mprgp_qp *q : qps) mprgp_qp_destroy(q);
  for (mprgp_matrix *A : mats) mprgp_matrix_destroy(A);
  mprgp_ctx_destroy(ctx);
  std::printf("graph_churn ok: %ld solves\n", solves);
  std::fflush(stdout);
  // Every object of the library is destroyed above, so the sanitizer has
  // seen each of its frees.  The process then ends without the HIP runtime's
  // own static destructors: under the sanitizer's device-memory support they
  // free runtime memory after the device runtime has shut down, which fails
  // the sanitizer's internal check (sanitizer_allocator_device.h) inside
  // libamdhip64/libhsa-runtime64, after this program's work is complete.
  _exit(0);
}
