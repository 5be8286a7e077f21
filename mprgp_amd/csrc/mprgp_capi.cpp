// mprgp_capi.cpp -- C ABI implementation: contexts, matrices, QPs and the
// batched device-resident solve loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <memory>
#include <numeric>
#include <thread>

#include "mprgp_host.h"
#include "mprgp_qpfile.h"

namespace mprgp {

// |v| of the power method's start vector (power_start_raw, summed in index
// order; the device divides each element by it).
double power_start_norm(size_t n) {
  double nr = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double e = power_start_raw(i);
    nr += e * e;
  }
  return std::sqrt(nr);
}

}  // namespace mprgp

using namespace mprgp;

namespace {

constexpr int kHostThreads = 16;

// Runs f(t, i0, i1) over [0, n) in contiguous ranges, range t on host thread
// t (range 0 on the caller's): up to kHostThreads, and none with fewer than
// `grain` items.  An exception in any range (or a thread that cannot start,
// whose range then runs on the caller's thread) surfaces as std::bad_alloc
// once every thread has joined.
template <class F>
void host_parallel(int n, int grain, F &&f) {
  const unsigned hw = std::thread::hardware_concurrency();
  int nt = (int)std::min(hw ? hw : 1u, (unsigned)kHostThreads);
  nt = std::max(1, std::min(nt, n / std::max(grain, 1)));
  auto bounds = [&](int t, int &i0, int &i1) {
    i0 = (int)((long long)n * t / nt);
    i1 = (int)((long long)n * (t + 1) / nt);
  };
  char failed[kHostThreads] = {0};
  auto run = [&](int t) {
    int i0, i1;
    bounds(t, i0, i1);
    try {
      f(t, i0, i1);
    } catch (...) {
      failed[t] = 1;
    }
  };
  std::vector<std::thread> th;
  int started = 1;
  try {
    th.reserve(nt);
    for (; started < nt; ++started) th.emplace_back(run, started);
  } catch (...) {
  }
  run(0);
  for (int t = started; t < nt; ++t) run(t);
  for (auto &x : th) x.join();
  for (int t = 0; t < nt; ++t)
    if (failed[t]) throw std::bad_alloc();
}

}  // namespace

int mprgp_ctx::alloc(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  std::lock_guard<std::mutex> lock(amu);
  if (budget && in_use + bytes > budget)
    return fail(MPRGP_EBUDGET, "HBM budget exceeded: %zu in use + %zu requested > %zu", in_use, bytes,
                budget);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(MPRGP_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  in_use += bytes;
  return MPRGP_OK;
}

void mprgp_ctx::release(void *p, size_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 256;
  std::lock_guard<std::mutex> lock(amu);
  (void)hipFree(p);
  in_use -= std::min(in_use, bytes);
}

int mprgp_ctx::scratch_reserve(size_t bytes) {
  if (bytes <= scratch_bytes) return MPRGP_OK;
  release(scratch, scratch_bytes);
  scratch = nullptr;
  scratch_bytes = 0;
  if (int rc = alloc(&scratch, bytes)) return rc;
  scratch_bytes = bytes;
  return MPRGP_OK;
}

// ------------------------------------------------------------------ context

extern "C" void mprgp_params_default(mprgp_params *p) {
  MPRGP_TRY
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->tol = 1e-3;  // MPRGPDecoupledCon::solve default
  p->max_it = 1000;
  p->variant = MPRGP_MONOTONIC;
  p->precision = MPRGP_FP64;
  p->preconditioned = 1;
  p->project_first = 0;
  p->use_ev = 1;
  p->alpha_bar = 0.0;
  p->power_tol = 0.0;
  MPRGP_CATCH_VOID
}

extern "C" int mprgp_device_count(int *count) {
  MPRGP_TRY
  if (!count) return fail(MPRGP_EBADINPUT, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return fail(MPRGP_ENODEV, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *count = n;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_ctx_create(int device, size_t hbm_budget_bytes, mprgp_ctx **out) {
  MPRGP_TRY
  if (!out) return fail(MPRGP_EBADINPUT, "out is NULL");
  *out = nullptr;
  int n = 0;
  if (mprgp_device_count(&n) != MPRGP_OK || n <= 0) return fail(MPRGP_ENODEV, "no HIP device available");
  if (device < 0 || device >= n) return fail(MPRGP_ENODEV, "device %d out of range (%d devices)", device, n);
  MPRGP_HIP(hipSetDevice(device));
  mprgp_ctx *c = new mprgp_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_running, 2 * sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&c->h_final, 2 * sizeof(QpState), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&c->t0);
  if (e == hipSuccess) e = hipEventCreate(&c->t1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->q[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->q[1], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ready, hipEventDisableTiming);
  if (e != hipSuccess) {
    mprgp_ctx_destroy(c);
    return fail(MPRGP_EHIP, "context setup failed: %s", hipGetErrorString(e));
  }
  int rc = c->tab.init(c);
  if (!rc) rc = noalloc_setup(device);
  if (rc) {
    mprgp_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" void mprgp_ctx_destroy(mprgp_ctx *c) {
  MPRGP_TRY
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->tab.destroy();
  c->release(c->scratch, c->scratch_bytes);
  if (c->h_running) (void)hipHostFree(c->h_running);
  if (c->h_final) (void)hipHostFree(c->h_final);
  if (c->t0) (void)hipEventDestroy(c->t0);
  if (c->t1) (void)hipEventDestroy(c->t1);
  if (c->q[0]) (void)hipEventDestroy(c->q[0]);
  if (c->q[1]) (void)hipEventDestroy(c->q[1]);
  if (c->ready) (void)hipEventDestroy(c->ready);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  MPRGP_CATCH_VOID
}

extern "C" int mprgp_ctx_memory(mprgp_ctx *c, size_t *in_use, size_t *budget) {
  MPRGP_TRY
  if (!c) return fail(MPRGP_EBADINPUT, "ctx is NULL");
  if (in_use) *in_use = c->in_use;
  if (budget) *budget = c->budget;
  return MPRGP_OK;
  MPRGP_CATCH
}

// ------------------------------------------------------------------ matrix

// Builds the SELL-64 layout from a node-order block CSR and uploads it.
int mprgp::build_matrix(mprgp_ctx *ctx, int n_nodes, const int *rowptr, const int *colidx,
                        const double *blocks, mprgp_matrix **out) {
  if (!ctx || !out || n_nodes <= 0 || !rowptr || !colidx)
    return fail(MPRGP_EBADINPUT, "bad matrix arguments");
  *out = nullptr;
  const long long nb = rowptr[n_nodes];
  if (rowptr[0] != 0 || nb <= 0) return fail(MPRGP_EBADINPUT, "rowptr must start at 0 and be non-empty");
  // Checked in the order a row-by-row pass would meet the faults: the first
  // row whose rowptr decreases; before it, the first row with a column out of
  // range or no diagonal block; then the first non-finite value.
  int bad_ptr = n_nodes;
  std::vector<int> len(n_nodes);
  for (int i = 0; i < n_nodes; ++i) {
    if (rowptr[i + 1] < rowptr[i]) {
      bad_ptr = i;
      break;
    }
    len[i] = rowptr[i + 1] - rowptr[i];
  }
  int bad_row[kHostThreads];
  std::fill(bad_row, bad_row + kHostThreads, INT_MAX);
  host_parallel(bad_ptr, 4096, [&](int t, int i0, int i1) {
    for (int i = i0; i < i1; ++i) {
      bool ok = false;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        if (colidx[k] < 0 || colidx[k] >= n_nodes) {
          ok = false;
          break;
        }
        ok |= colidx[k] == i;
      }
      if (!ok) {
        bad_row[t] = i;
        return;
      }
    }
  });
  const int first_bad = *std::min_element(bad_row, bad_row + kHostThreads);
  if (first_bad != INT_MAX) {
    for (int k = rowptr[first_bad]; k < rowptr[first_bad + 1]; ++k)
      if (colidx[k] < 0 || colidx[k] >= n_nodes)
        return fail(MPRGP_EBADINPUT, "column index %d out of range in row %d", colidx[k], first_bad);
    return fail(MPRGP_EBADINPUT, "row %d has no diagonal block", first_bad);
  }
  if (bad_ptr < n_nodes) return fail(MPRGP_EBADINPUT, "rowptr not monotone at row %d", bad_ptr);
  if (blocks) {
    long long bad_v[kHostThreads];
    std::fill(bad_v, bad_v + kHostThreads, LLONG_MAX);
    host_parallel(n_nodes, 4096, [&](int t, int i0, int i1) {
      for (long long k = 9LL * rowptr[i0]; k < 9LL * rowptr[i1]; ++k)
        if (!std::isfinite(blocks[k])) {
          bad_v[t] = k;
          return;
        }
    });
    const long long k = *std::min_element(bad_v, bad_v + kHostThreads);
    if (k != LLONG_MAX) return fail(MPRGP_EBADINPUT, "non-finite matrix value at %lld", k);
  }
  mprgp_matrix *A = new mprgp_matrix();
  A->ctx = ctx;
  A->n_nodes = n_nodes;
  A->nb = nb;
  A->perm.resize(n_nodes);
  std::iota(A->perm.begin(), A->perm.end(), 0);
  // internal order: descending block-row length, stable (keeps mesh locality
  // within each length class), over the whole matrix or, when
  // MPRGP_SELL_SIGMA is set, within windows of that many consecutive nodes
  {
    const char *e = getenv("MPRGP_SELL_SIGMA");
    const long long sigma = (e && e[0]) ? std::max(0LL, atoll(e)) : 0;
    const long long w = sigma > 0 ? sigma : (long long)n_nodes;
    for (long long a0 = 0; a0 < n_nodes; a0 += w) {
      const long long a1 = std::min<long long>(n_nodes, a0 + w);
      std::stable_sort(A->perm.begin() + a0, A->perm.begin() + a1, [&](int a, int b) { return len[a] > len[b]; });
    }
  }
  A->iperm.resize(n_nodes);
  for (int i = 0; i < n_nodes; ++i) A->iperm[A->perm[i]] = i;
  A->n_slices = (n_nodes + kSlice - 1) / kSlice;
  A->sl_off_h.resize(A->n_slices + 1);
  A->sl_off_h[0] = 0;
  long long slots = 0;
  for (int s = 0; s < A->n_slices; ++s) {
    int w = 0;
    for (int l = 0; l < kSlice; ++l) {
      int i = s * kSlice + l;
      if (i < n_nodes) w = std::max(w, len[A->perm[i]]);
    }
    slots += (long long)w * kSlice;
    if (slots > 0x7fffffffLL / 9) {
      delete A;
      return fail(MPRGP_EBADINPUT, "matrix too large for 32-bit SELL offsets");
    }
    A->sl_off_h[s + 1] = (int)slots;
  }
  A->slots = slots;
  // padding slots point at a valid node (the row itself, or the last node for
  // lanes past the end) with zero values, so the SpMV needs no branch; the
  // slices fill in parallel (each writes its own slots)
  std::unique_ptr<int[]> col;
  std::unique_ptr<double[]> val;
  try {
    col.reset(new int[slots]);
    val.reset(new double[(size_t)slots * 9]);
    A->blk_slot.resize(nb);
  } catch (...) {
    delete A;
    throw;
  }
  auto fill = [&](int, int s0, int s1) {
    for (int s = s0; s < s1; ++s) {
      const long long b0 = A->sl_off_h[s], b1 = A->sl_off_h[s + 1];
      std::memset(val.get() + b0 * 9, 0, sizeof(double) * 9 * (size_t)(b1 - b0));
      for (long long k = b0; k < b1; ++k) {
        const int i = s * kSlice + (int)((k - b0) % kSlice);
        col[k] = i < n_nodes ? i : n_nodes - 1;
      }
      for (int l = 0; l < kSlice; ++l) {
        const int i = s * kSlice + l;
        if (i >= n_nodes) continue;
        const int o = A->perm[i];
        for (int k = 0; k < len[o]; ++k) {
          const long long src = rowptr[o] + k;
          const long long slot = b0 + (long long)k * kSlice + l;
          col[slot] = A->iperm[colidx[src]];
          A->blk_slot[src] = slot;
          const long long vbase = b0 * 9 + (long long)k * 9 * kSlice + l;
          if (blocks)
            for (int e = 0; e < 9; ++e) val[vbase + (long long)e * kSlice] = blocks[9 * src + e];
        }
      }
    }
  };
  try {
    host_parallel(A->n_slices, 64, fill);
  } catch (...) {
    delete A;
    throw;
  }
  size_t b_off = sizeof(int) * (A->n_slices + 1), b_col = sizeof(int) * slots,
         b_val = sizeof(double) * slots * 9, b_d = sizeof(double) * 3 * n_nodes, b_p = sizeof(int) * n_nodes;
  // 16-bit columns for matrices whose node indices fit
  const bool c16 = n_nodes <= 65536;
  const size_t b_c16 = c16 ? sizeof(unsigned short) * slots : 0;
  int rc = ctx->alloc((void **)&A->d_sl_off, b_off);
  if (!rc) rc = ctx->alloc((void **)&A->d_col, b_col);
  if (!rc && c16) rc = ctx->alloc((void **)&A->d_col16, b_c16);
  if (!rc) rc = ctx->alloc((void **)&A->d_val, b_val);
  if (!rc) rc = ctx->alloc((void **)&A->d_dinv, b_d);
  if (!rc) rc = ctx->alloc((void **)&A->d_perm, b_p);
  A->bytes = b_off + b_col + b_c16 + b_val + b_d + b_p;
  if (rc) {
    mprgp_matrix_destroy(A);
    return rc;
  }
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(A->d_sl_off, A->sl_off_h.data(), b_off, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(A->d_col, col.get(), b_col, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && c16) e = launch_to_u16(s, A->d_col, A->d_col16, (size_t)slots);
  if (e == hipSuccess) e = hipMemcpyAsync(A->d_val, val.get(), b_val, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(A->d_perm, A->perm.data(), b_p, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_setup_diag(s, A->d_sl_off, A->d_col, A->d_val, A->d_dinv, n_nodes);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    mprgp_matrix_destroy(A);
    return fail(MPRGP_EHIP, "matrix upload failed: %s", hipGetErrorString(e));
  }
  *out = A;
  return MPRGP_OK;
}

extern "C" int mprgp_matrix_create_bsr3(mprgp_ctx *ctx, int n_nodes, const int *rowptr, const int *colidx,
                                        const double *blocks, mprgp_matrix **out) {
  MPRGP_TRY
  if (!ctx || !blocks) return fail(MPRGP_EBADINPUT, "ctx or blocks is NULL");
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  return build_matrix(ctx, n_nodes, rowptr, colidx, blocks, out);
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_create_csr(mprgp_ctx *ctx, int n, const int *ptr, const int *idx,
                                       const double *val, mprgp_matrix **out) {
  MPRGP_TRY
  if (!ctx || !ptr || !idx || !val || !out) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (n <= 0 || n % 3) return fail(MPRGP_EBADINPUT, "n = %d is not a positive multiple of 3", n);
  const int N = n / 3;
  // Checked in the order a row-by-row pass would meet the faults: the first
  // scalar row whose ptr decreases, and before it the first column index out
  // of range.
  int bad_row = n;
  for (int row = 0; row < n; ++row)
    if (ptr[row + 1] < ptr[row]) {
      bad_row = row;
      break;
    }
  long long bad_k[kHostThreads];
  std::fill(bad_k, bad_k + kHostThreads, LLONG_MAX);
  host_parallel(bad_row, 8192, [&](int t, int r0, int r1) {
    for (int row = r0; row < r1; ++row)
      for (int k = ptr[row]; k < ptr[row + 1]; ++k)
        if (idx[k] < 0 || idx[k] >= n) {
          bad_k[t] = k;
          return;
        }
  });
  const long long first_bad = *std::min_element(bad_k, bad_k + kHostThreads);
  if (first_bad != LLONG_MAX) return fail(MPRGP_EBADINPUT, "index %d out of range", idx[first_bad]);
  if (bad_row < n) return fail(MPRGP_EBADINPUT, "ptr not monotone at %d", bad_row);
  // node-row blocks: the distinct block columns of the 3 scalar rows plus the
  // (possibly zero) diagonal block, ascending; scalars summed into them in
  // CSR order.  Counted first, then filled, node rows in parallel.
  std::vector<int> rowptr(N + 1, 0);
  host_parallel(N, 2048, [&](int, int a0, int a1) {
    std::vector<int> mark(N, -1);
    for (int a = a0; a < a1; ++a) {
      int c = 0;
      for (int k = ptr[3 * a]; k < ptr[3 * a + 3]; ++k) {
        const int b = idx[k] / 3;
        if (mark[b] != a) {
          mark[b] = a;
          ++c;
        }
      }
      rowptr[a + 1] = c + (mark[a] != a ? 1 : 0);
    }
  });
  for (int a = 0; a < N; ++a) rowptr[a + 1] += rowptr[a];
  std::vector<int> colidx(rowptr[N]);
  // zeroed row by row in the parallel fill (first touch on the filling thread)
  std::unique_ptr<double[]> blocks(new double[(size_t)rowptr[N] * 9]);
  host_parallel(N, 2048, [&](int, int a0, int a1) {
    std::vector<int> mark(N, -1), slot(N, -1), cols;
    for (int a = a0; a < a1; ++a) {
      cols.clear();
      for (int k = ptr[3 * a]; k < ptr[3 * a + 3]; ++k) {
        const int b = idx[k] / 3;
        if (mark[b] != a) {
          mark[b] = a;
          cols.push_back(b);
        }
      }
      if (mark[a] != a) cols.push_back(a);  // keep a (zero) diagonal block
      std::sort(cols.begin(), cols.end());
      const int base = rowptr[a];
      std::memset(blocks.get() + (size_t)base * 9, 0, sizeof(double) * 9 * cols.size());
      for (size_t k = 0; k < cols.size(); ++k) {
        slot[cols[k]] = base + (int)k;
        colidx[base + k] = cols[k];
      }
      for (int r = 0; r < 3; ++r)
        for (int k = ptr[3 * a + r]; k < ptr[3 * a + r + 1]; ++k)
          blocks[(size_t)slot[idx[k] / 3] * 9 + r * 3 + idx[k] % 3] += val[k];
    }
  });
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  return build_matrix(ctx, N, rowptr.data(), colidx.data(), blocks.get(), out);
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_update_bsr3(mprgp_matrix *A, const double *blocks) {
  MPRGP_TRY
  if (!A || !blocks) return fail(MPRGP_EBADINPUT, "NULL argument");
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  // the first non-finite value in block order; nothing changes if there is one
  const int nb = (int)A->nb;
  long long bad_v[kHostThreads];
  std::fill(bad_v, bad_v + kHostThreads, LLONG_MAX);
  host_parallel(nb, 16384, [&](int t, int b0, int b1) {
    for (long long k = 9LL * b0; k < 9LL * b1; ++k)
      if (!std::isfinite(blocks[k])) {
        bad_v[t] = k;
        return;
      }
  });
  const long long bad = *std::min_element(bad_v, bad_v + kHostThreads);
  if (bad != LLONG_MAX) return fail(MPRGP_EBADINPUT, "non-finite matrix value at %lld", bad);
  // SELL values: every slice zeroes its slots (the padding stays zero), then
  // each block goes to its slot, both on the host threads
  std::unique_ptr<double[]> val(new double[(size_t)A->slots * 9]);
  host_parallel(A->n_slices, 64, [&](int, int s0, int s1) {
    const long long a0 = A->sl_off_h[s0], a1 = A->sl_off_h[s1];
    std::memset(val.get() + a0 * 9, 0, sizeof(double) * 9 * (size_t)(a1 - a0));
  });
  host_parallel(nb, 16384, [&](int, int b0, int b1) {
    for (int b = b0; b < b1; ++b) {
      const long long slot = A->blk_slot[b];
      // slot = s0 + 64 k + l for the start s0 of its slice (a multiple of
      // 64): the values sit at 9 s0 + 576 k + l + 64 e
      const long long l = slot % kSlice;
      const int s = (int)(std::upper_bound(A->sl_off_h.begin(), A->sl_off_h.end(), (int)slot) -
                          A->sl_off_h.begin()) - 1;
      const long long s0 = A->sl_off_h[s];
      const long long k = (slot - l - s0) / kSlice;
      const long long vbase = s0 * 9 + k * 9 * kSlice + l;
      for (int e = 0; e < 9; ++e) val[vbase + (long long)e * kSlice] = blocks[9LL * b + e];
    }
  });
  hipStream_t s = ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(A->d_val, val.get(), sizeof(double) * 9 * (size_t)A->slots, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_setup_diag(s, A->d_sl_off, A->d_col, A->d_val, A->d_dinv, A->n_nodes));
  MPRGP_HIP(hipStreamSynchronize(s));
  A->version++;
  A->lambda_min_lower = 0.0;  // new values: the old bound no longer holds
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_get_bsr3(mprgp_matrix *A, int *rowptr, int *colidx, double *blocks) {
  MPRGP_TRY
  if (!A || !rowptr || !colidx || !blocks) return fail(MPRGP_EBADINPUT, "NULL argument");
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  std::vector<int> col((size_t)A->slots);
  std::vector<double> val((size_t)A->slots * 9);
  hipStream_t s = ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(col.data(), A->d_col, sizeof(int) * col.size(), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipMemcpyAsync(val.data(), A->d_val, sizeof(double) * val.size(), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  // block b of the creation order sits in SELL slot blk_slot[b]: its row is
  // the slot's lane in its slice, its column the slot's column index, both
  // internal node numbers; creation order is by row, so the row counts give
  // rowptr
  std::fill(rowptr, rowptr + A->n_nodes + 1, 0);
  for (long long b = 0; b < A->nb; ++b) {
    const long long slot = A->blk_slot[b];
    const int sl = (int)(std::upper_bound(A->sl_off_h.begin(), A->sl_off_h.end(), (int)slot) -
                         A->sl_off_h.begin()) - 1;
    const long long b0 = A->sl_off_h[sl];
    const long long l = (slot - b0) % kSlice, k = (slot - b0) / kSlice;
    const int row = A->perm[sl * kSlice + (int)l];
    rowptr[row + 1]++;
    colidx[b] = A->perm[col[slot]];
    const long long vbase = b0 * 9 + k * 9 * kSlice + l;
    for (int e = 0; e < 9; ++e) blocks[9 * b + e] = val[vbase + (long long)e * kSlice];
  }
  for (int i = 0; i < A->n_nodes; ++i) rowptr[i + 1] += rowptr[i];
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_set_lambda_min(mprgp_matrix *A, double lower) {
  MPRGP_TRY
  if (!A) return fail(MPRGP_EBADINPUT, "NULL matrix");
  if (!(lower >= 0.0) || !std::isfinite(lower)) return fail(MPRGP_EBADINPUT, "bound must be finite and >= 0");
  A->lambda_min_lower = lower;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_info(const mprgp_matrix *A, int *n_nodes, long long *n_blocks, long long *bytes) {
  MPRGP_TRY
  if (!A) return fail(MPRGP_EBADINPUT, "NULL matrix");
  if (n_nodes) *n_nodes = A->n_nodes;
  if (n_blocks) *n_blocks = A->nb;
  if (bytes) *bytes = (long long)A->bytes;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_spmv(mprgp_matrix *A, const double *x, double *y) {
  MPRGP_TRY
  if (!A || !x || !y) return fail(MPRGP_EBADINPUT, "NULL argument");
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  const int n = A->n_nodes;
  const size_t vb = sizeof(double) * 3 * n;
  double *buf = nullptr;
  int rc = ctx->alloc((void **)&buf, 3 * vb);
  if (rc) return rc;
  double *du = buf, *dx = buf + 3 * n, *dy = buf + 6 * n;
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemcpyAsync(du, x, vb, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = launch_gather_nodes(s, du, dx, A->d_perm, n);
  if (e == hipSuccess) e = launch_spmv_plain(s, A->d_sl_off, A->d_col, A->d_val, dx, dy, n, 0);
  if (e == hipSuccess) e = launch_scatter_nodes(s, dy, du, A->d_perm, n);
  if (e == hipSuccess) e = hipMemcpyAsync(y, du, vb, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  ctx->release(buf, 3 * vb);
  if (e != hipSuccess) return fail(MPRGP_EHIP, "spmv failed: %s", hipGetErrorString(e));
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_diagonal(mprgp_matrix *A, double *diag) {
  MPRGP_TRY
  if (!A || !diag) return fail(MPRGP_EBADINPUT, "NULL argument");
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  const int n = A->n_nodes;
  const size_t vb = sizeof(double) * 3 * n;
  if (int rc = ctx->scratch_reserve(2 * vb + 256)) return rc;
  double *dd = (double *)ctx->scratch, *du = dd + 3 * (size_t)n;
  hipStream_t s = ctx->stream;
  MPRGP_HIP(launch_get_diag(s, A->d_sl_off, A->d_col, A->d_val, dd, n));
  MPRGP_HIP(launch_scatter_nodes(s, dd, du, A->d_perm, n));
  MPRGP_HIP(hipMemcpyAsync(diag, du, vb, hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_matrix_func_value(mprgp_matrix *A, const double *x, const double *b, double *f) {
  MPRGP_TRY
  if (!A || !x || !b || !f) return fail(MPRGP_EBADINPUT, "NULL argument");
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  const int n = A->n_nodes;
  const size_t N3 = 3 * (size_t)n;
  const size_t blocks = (N3 + 255) / 256;
  if (int rc = ctx->scratch_reserve(sizeof(double) * (4 * N3 + blocks + 2) + 256)) return rc;
  double *du = (double *)ctx->scratch, *dx = du + N3, *db = dx + N3, *dy = db + N3;
  double *part = dy + N3, *res = part + blocks;
  hipStream_t s = ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(du, x, sizeof(double) * N3, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_gather_nodes(s, du, dx, A->d_perm, n));
  MPRGP_HIP(hipMemcpyAsync(du, b, sizeof(double) * N3, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_gather_nodes(s, du, db, A->d_perm, n));
  MPRGP_HIP(launch_spmv_plain(s, A->d_sl_off, A->d_col, A->d_val, dx, dy, n, 0));
  // 0.5 x.(A x) - x.b  (FixedSparseMatrix::funcValue, MPRGPSolver.h:475-482)
  MPRGP_HIP(launch_dot(s, dx, dy, N3, part, res));
  MPRGP_HIP(launch_dot(s, dx, db, N3, part, res + 1));
  double r[2] = {0.0, 0.0};
  MPRGP_HIP(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  *f = 0.5 * r[0] - r[1];
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" void mprgp_matrix_destroy(mprgp_matrix *A) {
  MPRGP_TRY
  if (!A) return;
  mprgp_ctx *ctx = A->ctx;
  (void)hipSetDevice(ctx->device);
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->release(A->d_sl_off, sizeof(int) * (A->n_slices + 1));
  ctx->release(A->d_col, sizeof(int) * A->slots);
  ctx->release(A->d_col16, sizeof(unsigned short) * A->slots);
  ctx->release(A->d_val, sizeof(double) * A->slots * 9);
  ctx->release(A->d_dinv, sizeof(double) * 3 * A->n_nodes);
  ctx->release(A->d_perm, sizeof(int) * A->n_nodes);
  ctx->release(A->d_val32, sizeof(float) * A->slots * 9);
  delete A;
  MPRGP_CATCH_VOID
}

// FP32 copy of the values for the mixed form.
static int ensure_matrix32(mprgp_matrix *A, hipStream_t s) {
  mprgp_ctx *ctx = A->ctx;
  if (!A->d_val32) {
    if (int rc = ctx->alloc((void **)&A->d_val32, sizeof(float) * A->slots * 9)) return rc;
  }
  if (A->version32 != A->version) {
    MPRGP_HIP(launch_to_f32(s, A->d_val, A->d_val32, (size_t)A->slots * 9));
    A->version32 = A->version;
  }
  return MPRGP_OK;
}

// s: the stream the QP's solve will run on (the engine's own, for a QP
// submitted to an engine), so the clearing is ordered before the solve.
static int ensure_qp32(mprgp_qp *q, hipStream_t s) {
  if (q->base32) return MPRGP_OK;
  const size_t n3 = 3 * (size_t)q->n;
  const size_t bytes = (3 * n3 + (size_t)q->n) * sizeof(double) + (9 * n3 + 64) * sizeof(float);
  if (int rc = q->ctx->alloc(&q->base32, bytes)) return rc;
  q->bytes32 = bytes;
  double *dp = (double *)q->base32;
  double **dvecs[] = {&q->fx, &q->fy, &q->fpy};
  for (double **v : dvecs) {
    *v = dp;
    dp += n3;
  }
  q->fcc = dp;
  dp += q->n;
  float *fp32 = (float *)dp;
  float **fvecs[] = {&q->fp, &q->fp2, &q->fz, &q->fap, &q->fb, &q->fg, &q->fphi, &q->fbeta, &q->fbeta2};
  for (float **v : fvecs) {
    *v = fp32;
    fp32 += n3;
  }
  MPRGP_HIP(hipMemsetAsync(q->base32, 0, q->bytes32, s));
  return MPRGP_OK;
}

// ---------------------------------------------------------------------- QP

static size_t qp_tiles_max(int n_nodes) { return (size_t)((n_nodes + kSlice - 1) / kSlice); }

extern "C" int mprgp_qp_create(mprgp_matrix *A, mprgp_qp **out) {
  MPRGP_TRY
  if (!A || !out) return fail(MPRGP_EBADINPUT, "NULL argument");
  *out = nullptr;
  mprgp_ctx *ctx = A->ctx;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  mprgp_qp *q = new mprgp_qp();
  q->ctx = ctx;
  q->A = A;
  q->n = A->n_nodes;
  const size_t n3 = 3 * (size_t)q->n;
  const int n_vec = 19;      // x x0 b g p z phi beta y ap py ev pv pw stage p2 beta2 pu pz
  const int n_row = 3 + 3;   // cn(3) cc cj cs
  q->part_cap = (int)qp_tiles_max(q->n);
  size_t doubles = n_vec * n3 + n_row * (size_t)q->n + 2 * (size_t)q->part_cap * kNumPart;
  size_t bytes = doubles * sizeof(double) + 2 * sizeof(QpState) + 2 * (size_t)q->n + 512;
  int rc = ctx->alloc((void **)&q->base, bytes);
  if (rc) {
    delete q;
    return rc;
  }
  q->bytes = bytes;
  double *p = q->base;
  double **vecs[] = {&q->x, &q->x0, &q->b, &q->g, &q->p, &q->z, &q->phi, &q->beta,
                     &q->y, &q->ap, &q->py, &q->ev, &q->pv, &q->pw, &q->stage, &q->p2, &q->beta2,
                     &q->pu, &q->pz};
  for (double **v : vecs) {
    *v = p;
    p += n3;
  }
  q->cn = p; p += n3;
  q->cc = p; p += q->n;
  q->cj = p; p += q->n;
  q->cs = p; p += q->n;
  q->part = p; p += 2 * (size_t)q->part_cap * kNumPart;
  uintptr_t u = (uintptr_t)p;
  u = (u + 127) & ~(uintptr_t)127;
  q->st = (QpState *)u;
  q->cf = (unsigned char *)(q->st + 2);
  q->face = q->cf + q->n;
  hipStream_t s = ctx->stream;
  hipError_t e = hipMemsetAsync(q->base, 0, bytes, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    mprgp_qp_destroy(q);
    return fail(MPRGP_EHIP, "qp init failed: %s", hipGetErrorString(e));
  }
  *out = q;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" void mprgp_qp_destroy(mprgp_qp *q) {
  MPRGP_TRY
  if (!q) return;
  (void)hipSetDevice(q->ctx->device);
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  (void)hipStreamSynchronize(q->ctx->stream);
  q->ctx->release(q->base, q->bytes);
  q->ctx->release(q->base32, q->bytes32);
  delete q;
  MPRGP_CATCH_VOID
}

// host node-order vector -> device internal order
static int upload_vec(mprgp_qp *q, const double *h, double *dst) {
  const size_t vb = sizeof(double) * 3 * q->n;
  hipStream_t s = q->ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(q->stage, h, vb, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_gather_nodes(s, q->stage, dst, q->A->d_perm, q->n));
  MPRGP_HIP(hipStreamSynchronize(s));
  return MPRGP_OK;
}
static int download_vec(mprgp_qp *q, const double *src, double *h) {
  const size_t vb = sizeof(double) * 3 * q->n;
  hipStream_t s = q->ctx->stream;
  MPRGP_HIP(launch_scatter_nodes(s, src, q->stage, q->A->d_perm, q->n));
  MPRGP_HIP(hipMemcpyAsync(h, q->stage, vb, hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  return MPRGP_OK;
}

static int check_finite(const double *v, size_t n, const char *what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return fail(MPRGP_EBADINPUT, "%s[%zu] is not finite", what, i);
  return MPRGP_OK;
}

extern "C" int mprgp_qp_set_rhs(mprgp_qp *q, const double *b) {
  MPRGP_TRY
  if (!q || !b) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (int rc = check_finite(b, 3 * (size_t)q->n, "b")) return rc;
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  int rc = upload_vec(q, b, q->b);
  if (!rc) q->has_rhs = true;
  return rc;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_set_x(mprgp_qp *q, const double *x) {
  MPRGP_TRY
  if (!q || !x) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (int rc = check_finite(x, 3 * (size_t)q->n, "x")) return rc;
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  int rc = upload_vec(q, x, q->x0);
  if (rc) return rc;
  MPRGP_HIP(hipMemcpyAsync(q->x, q->x0, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, q->ctx->stream));
  MPRGP_HIP(hipStreamSynchronize(q->ctx->stream));
  q->has_x = true;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_reset_x(mprgp_qp *q) {
  MPRGP_TRY
  if (!q) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!q->has_x) return fail(MPRGP_ESTATE, "x was never set");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  // queued on the context's stream; a solve on another stream (an engine's)
  // is ordered after it by prepare_qp
  MPRGP_HIP(hipMemcpyAsync(q->x, q->x0, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, q->ctx->stream));
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_get_x(mprgp_qp *q, double *x) {
  MPRGP_TRY
  if (!q || !x) return fail(MPRGP_EBADINPUT, "NULL argument");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  return download_vec(q, q->x, x);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_get_rhs(mprgp_qp *q, double *b) {
  MPRGP_TRY
  if (!q || !b) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!q->has_rhs) return fail(MPRGP_ESTATE, "rhs was never set");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  return download_vec(q, q->b, b);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_set_ev(mprgp_qp *q, const double *ev) {
  MPRGP_TRY
  if (!q) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!ev) {
    q->ev_valid = 0;
    return MPRGP_OK;
  }
  if (int rc = check_finite(ev, 3 * (size_t)q->n, "ev")) return rc;
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  int rc = upload_vec(q, ev, q->ev);
  if (!rc) q->ev_valid = 1;
  return rc;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_get_ev(mprgp_qp *q, double *ev, int *valid) {
  MPRGP_TRY
  if (!q) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (valid) *valid = q->ev_valid;
  if (!ev || !q->ev_valid) return MPRGP_OK;
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  return download_vec(q, q->ev, ev);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_op_counts(mprgp_qp *q, int *counts, int n) {
  MPRGP_TRY
  if (!q || !counts || n <= 0) return fail(MPRGP_EBADINPUT, "bad arguments");
  for (int i = 0; i < n; ++i) counts[i] = i < OP_NUM ? q->opcnt[i] : 0;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_get_face(mprgp_qp *q, unsigned char *face) {
  MPRGP_TRY
  if (!q || !face) return fail(MPRGP_EBADINPUT, "NULL argument");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  std::vector<unsigned char> f(q->n);
  MPRGP_HIP(hipMemcpyAsync(f.data(), q->face, q->n, hipMemcpyDeviceToHost, q->ctx->stream));
  MPRGP_HIP(hipStreamSynchronize(q->ctx->stream));
  for (int i = 0; i < q->n; ++i) face[q->A->perm[i]] = f[i];
  return MPRGP_OK;
  MPRGP_CATCH
}

static int upload_rows(mprgp_qp *q, int m, const std::vector<int> &node, const std::vector<double> &nrm,
                       const std::vector<double> &c) {
  const int N = q->n;
  std::vector<double> cn(3 * (size_t)N, 0.0), cc(N, 0.0), cj(N, 0.0);
  std::vector<unsigned char> cf(N, 0);
  for (int i = 0; i < m; ++i) {
    int k = node[i];
    if (k < 0 || k >= N) return fail(MPRGP_EBADINPUT, "row %d: node %d out of range", i, k);
    int r = q->A->iperm[k];
    if (cf[r])
      return fail(MPRGP_EBADINPUT, "row %d: node %d already has a row (one half-space row per node is supported)",
                  i, k);
    double n0 = nrm[3 * i], n1 = nrm[3 * i + 1], n2 = nrm[3 * i + 2];
    if (!std::isfinite(n0) || !std::isfinite(n1) || !std::isfinite(n2) || !std::isfinite(c[i]))
      return fail(MPRGP_EBADINPUT, "row %d: non-finite normal or offset", i);
    double jj = n0 * n0 + n1 * n1 + n2 * n2;
    if (!(jj >= kEps)) return fail(MPRGP_EBADINPUT, "row %d: |n|^2 = %g below 1e-9", i, jj);
    cf[r] = 1;
    cn[3 * r] = n0;
    cn[3 * r + 1] = n1;
    cn[3 * r + 2] = n2;
    cc[r] = c[i];
    cj[r] = jj;
  }
  hipStream_t s = q->ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(q->cn, cn.data(), cn.size() * sizeof(double), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(q->cc, cc.data(), cc.size() * sizeof(double), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(q->cj, cj.data(), cj.size() * sizeof(double), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(q->cf, cf.data(), cf.size(), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  q->m = m;
  return MPRGP_OK;
}

extern "C" int mprgp_qp_set_rows(mprgp_qp *q, int m, const int *node, const double *normal, const double *c) {
  MPRGP_TRY
  if (!q || m < 0 || (m > 0 && (!node || !normal || !c))) return fail(MPRGP_EBADINPUT, "bad row arguments");
  if (m > q->n) return fail(MPRGP_EBADINPUT, "m = %d exceeds node count %d", m, q->n);
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  std::vector<int> nd(node, node + m);
  std::vector<double> nr(normal, normal + 3 * (size_t)m), cv(c, c + m);
  return upload_rows(q, m, nd, nr, cv);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_set_rows_sparse(mprgp_qp *q, int m, int col_major, const int *outer, const int *inner,
                                        const double *val, const double *c) {
  MPRGP_TRY
  if (!q || m < 0 || !outer || (m > 0 && (!inner || !val || !c)))
    return fail(MPRGP_EBADINPUT, "bad sparse row arguments");
  const int ncol = 3 * q->n;
  std::vector<int> node(m, -1);
  std::vector<double> nrm(3 * (size_t)m, 0.0);
  std::vector<double> cv(c, c + m);
  auto put = [&](int row, int colj, double v) -> int {
    if (row < 0 || row >= m || colj < 0 || colj >= ncol)
      return fail(MPRGP_EBADINPUT, "J entry (%d,%d) out of range", row, colj);
    int k = colj / 3;
    if (node[row] >= 0 && node[row] != k)
      return fail(MPRGP_EBADINPUT, "J row %d spans nodes %d and %d (only per-node rows are decoupled here)",
                  row, node[row], k);
    node[row] = k;
    nrm[3 * (size_t)row + colj % 3] += v;
    return MPRGP_OK;
  };
  if (col_major) {
    for (int j = 0; j < ncol; ++j)
      for (int k = outer[j]; k < outer[j + 1]; ++k)
        if (int rc = put(inner[k], j, val[k])) return rc;
  } else {
    for (int i = 0; i < m; ++i)
      for (int k = outer[i]; k < outer[i + 1]; ++k)
        if (int rc = put(i, inner[k], val[k])) return rc;
  }
  for (int i = 0; i < m; ++i)
    if (node[i] < 0) return fail(MPRGP_EBADINPUT, "J row %d is empty", i);
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  return upload_rows(q, m, node, nrm, cv);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_collide_planes(mprgp_qp *q, int n_planes, const double *planes, int *m_out) {
  MPRGP_TRY
  if (!q || n_planes < 0 || (n_planes > 0 && !planes)) return fail(MPRGP_EBADINPUT, "bad plane arguments");
  for (int k = 0; k < n_planes; ++k) {
    const double *pl = planes + 5 * (size_t)k;
    if (!std::isfinite(pl[0]) || !std::isfinite(pl[1]) || !std::isfinite(pl[2]) || !std::isfinite(pl[3]) ||
        std::isnan(pl[4]))
      return fail(MPRGP_EBADINPUT, "plane %d: non-finite normal or offset, or NaN band", k);
    const double nn = std::sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
    if (!(nn >= 1.0 - 1e-8 && nn <= 1.0 + 1e-8))  // LinearConCollider.cpp:145
      return fail(MPRGP_EBADINPUT, "plane %d: |n| = %.17g is not 1 within 1e-8", k, nn);
  }
  if (!q->has_x) return fail(MPRGP_ESTATE, "x was never set");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  mprgp_ctx *ctx = q->ctx;
  hipStream_t s = ctx->stream;
  // scratch: the row counter, then the planes from byte 256
  const size_t pbytes = 5 * sizeof(double) * (size_t)n_planes;
  if (int rc = ctx->scratch_reserve(256 + pbytes)) return rc;
  int *d_count = (int *)ctx->scratch;
  double *d_planes = (double *)((char *)ctx->scratch + 256);
  MPRGP_HIP(hipMemsetAsync(d_count, 0, sizeof(int), s));
  if (pbytes) MPRGP_HIP(hipMemcpyAsync(d_planes, planes, pbytes, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_collide_planes(s, q->x, d_planes, n_planes, q->cn, q->cc, q->cj, q->cf, d_count, q->n));
  int m = 0;
  MPRGP_HIP(hipMemcpyAsync(ctx->h_running, d_count, sizeof(int), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  m = ctx->h_running[0];
  q->m = m;
  if (m_out) *m_out = m;
  return MPRGP_OK;
  MPRGP_CATCH
}

// ------------------------------------------- projector and preconditioner

// One projector_op over host vectors in node order: a (and b) in, out and/or
// the reduction result back.  The face is the QP's own (internal order).
static int projector_call(mprgp_qp *q, int op, const double *a, const double *b, double *out, double s0,
                          double *result) {
  if (!q || !a) return fail(MPRGP_EBADINPUT, "NULL argument");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  mprgp_ctx *ctx = q->ctx;
  hipStream_t s = ctx->stream;
  const int N = q->n;
  const size_t n = 3 * (size_t)N;
  const int blocks = (N + 255) / 256;
  if (int rc = ctx->scratch_reserve(sizeof(double) * (5 * n + blocks + 1) + 1024)) return rc;
  double *ua = (double *)ctx->scratch, *ub = ua + n, *ia = ub + n, *ib = ia + n, *io = ib + n;
  double *part = io + n, *res = part + blocks;
  MPRGP_HIP(hipMemcpyAsync(ua, a, sizeof(double) * n, hipMemcpyHostToDevice, s));
  MPRGP_HIP(launch_gather_nodes(s, ua, ia, q->A->d_perm, N));
  if (b) {
    MPRGP_HIP(hipMemcpyAsync(ub, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
    MPRGP_HIP(launch_gather_nodes(s, ub, ib, q->A->d_perm, N));
  }
  if (op == PJ_PRECOND) MPRGP_HIP(launch_setup_rows(s, q->A->d_dinv, q->cn, q->cf, q->cs, N));
  ProjArgs g{};
  g.op = op;
  g.n_nodes = N;
  g.cf = q->cf;
  g.cn = q->cn;
  g.cc = q->cc;
  g.cj = q->cj;
  g.cs = q->cs;
  g.dinv = q->A->d_dinv;
  g.face = q->face;
  g.a = ia;
  g.b = b ? ib : ia;
  g.out = io;
  g.s0 = s0;
  g.part = result ? part : nullptr;
  MPRGP_HIP(launch_projector(s, g, res));
  if (out) {
    MPRGP_HIP(launch_scatter_nodes(s, io, ua, q->A->d_perm, N));
    MPRGP_HIP(hipMemcpyAsync(out, ua, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  }
  double r = 0.0;
  if (result) MPRGP_HIP(hipMemcpyAsync(&r, res, sizeof(double), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  if (result) *result = N > 0 ? r : (op == PJ_STEP_LIMIT ? kDblMax : 0.0);
  return MPRGP_OK;
}

extern "C" int mprgp_qp_project(mprgp_qp *q, const double *x, double *y) {
  MPRGP_TRY
  if (!y) return fail(MPRGP_EBADINPUT, "NULL argument");
  return projector_call(q, PJ_PROJECT, x, nullptr, y, 0.0, nullptr);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_is_feasible(mprgp_qp *q, const double *x, int *feasible) {
  MPRGP_TRY
  if (!feasible) return fail(MPRGP_EBADINPUT, "NULL argument");
  double violated = 0.0;
  if (int rc = projector_call(q, PJ_FEASIBLE, x, nullptr, nullptr, 0.0, &violated)) return rc;
  *feasible = violated == 0.0;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_decide_face(mprgp_qp *q, const double *x, int *active) {
  MPRGP_TRY
  double count = 0.0;
  if (int rc = projector_call(q, PJ_FACE, x, nullptr, nullptr, 0.0, &count)) return rc;
  if (active) *active = (int)count;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_phi(mprgp_qp *q, const double *g, double *phi) {
  MPRGP_TRY
  if (!phi) return fail(MPRGP_EBADINPUT, "NULL argument");
  return projector_call(q, PJ_PHI, g, nullptr, phi, 0.0, nullptr);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_beta(mprgp_qp *q, const double *g, const double *phi, double *beta) {
  MPRGP_TRY
  if (!phi || !beta) return fail(MPRGP_EBADINPUT, "NULL argument");
  return projector_call(q, PJ_BETA, g, phi, beta, 0.0, nullptr);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_phitphi(mprgp_qp *q, const double *x, double alpha_bar, const double *phi, double *out) {
  MPRGP_TRY
  if (!phi || !out) return fail(MPRGP_EBADINPUT, "NULL argument");
  if (!(alpha_bar > kEps)) return fail(MPRGP_EBADINPUT, "alpha_bar must exceed 1e-9");  // :387 assertion
  double s = 0.0;
  if (int rc = projector_call(q, PJ_PHITPHI, x, phi, nullptr, alpha_bar, &s)) return rc;
  const double r = s * (1.0 / alpha_bar);
  *out = r > 0.0 ? r : 0.0;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_step_limit(mprgp_qp *q, const double *x, const double *d, double alpha_cg,
                                   double *alpha_f) {
  MPRGP_TRY
  if (!d || !alpha_f) return fail(MPRGP_EBADINPUT, "NULL argument");
  double lim = kDblMax;
  if (int rc = projector_call(q, PJ_STEP_LIMIT, x, d, nullptr, 0.0, &lim)) return rc;
  double af = (alpha_cg == kDblMax) ? alpha_cg : (alpha_cg + kEps);
  *alpha_f = lim < af ? lim : af;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_qp_precond(mprgp_qp *q, const double *g, const double *phi, double *z) {
  MPRGP_TRY
  if (!phi || !z) return fail(MPRGP_EBADINPUT, "NULL argument");
  return projector_call(q, PJ_PRECOND, g, phi, z, q && q->m > 0 ? 1.0 : 0.0, nullptr);
  MPRGP_CATCH
}

extern "C" int mprgp_qp_get_rows(mprgp_qp *q, int cap, int *m, int *node, double *normal, double *c) {
  MPRGP_TRY
  if (!q || !m || cap < 0) return fail(MPRGP_EBADINPUT, "bad arguments");
  *m = q->m;
  if (!node && !normal && !c) return MPRGP_OK;  // count query
  if (q->m > cap) return fail(MPRGP_EBADINPUT, "the QP has %d rows, the arrays hold %d", q->m, cap);
  if (q->m > 0 && (!node || !normal || !c)) return fail(MPRGP_EBADINPUT, "NULL argument");
  MPRGP_NOT_BUSY(q);
  MPRGP_HIP(hipSetDevice(q->ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(q->ctx->mu);
  const int N = q->n;
  std::vector<double> cn(3 * (size_t)N), cc(N);
  std::vector<unsigned char> cf(N);
  hipStream_t s = q->ctx->stream;
  MPRGP_HIP(hipMemcpyAsync(cn.data(), q->cn, cn.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipMemcpyAsync(cc.data(), q->cc, cc.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipMemcpyAsync(cf.data(), q->cf, cf.size(), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  int i = 0;
  for (int k = 0; k < N; ++k) {
    const int r = q->A->iperm[k];
    if (!cf[r]) continue;
    if (i >= q->m) return fail(MPRGP_ESTATE, "row storage holds more than %d rows", q->m);
    node[i] = k;
    normal[3 * (size_t)i + 0] = cn[3 * (size_t)r + 0];
    normal[3 * (size_t)i + 1] = cn[3 * (size_t)r + 1];
    normal[3 * (size_t)i + 2] = cn[3 * (size_t)r + 2];
    c[i] = cc[r];
    ++i;
  }
  if (i != q->m) return fail(MPRGP_ESTATE, "row storage holds %d rows, expected %d", i, q->m);
  return MPRGP_OK;
  MPRGP_CATCH
}

// ----------------------------------------------------------- batch solve

namespace mprgp {

int BatchTables::init(mprgp_ctx *ctx) {
  this->ctx = ctx;
  if (int rc = ctx->alloc((void **)&d_ctrl, 4 * sizeof(int))) return rc;
  return MPRGP_OK;
}

void BatchTables::drop_graphs() {
  for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto &kv : graph_defs) (void)hipGraphDestroy(kv.second);
  graphs.clear();
  graph_defs.clear();
}

void BatchTables::destroy() {
  if (!ctx) return;
  drop_graphs();
  ctx->release(d_qps, sizeof(QpDev) * (size_t)qps_cap);
  ctx->release(d_tiles, sizeof(Tile) * (size_t)tiles_cap);
  if (h_tiles) (void)hipHostFree(h_tiles);
  ctx->release(d_ctrl, 4 * sizeof(int));
  ctx->release(d_flags, sizeof(int) * (size_t)flags_cap);
  if (h_flags) (void)hipHostFree(h_flags);
  ctx->release(d_states, 2 * sizeof(QpState) * (size_t)states_cap);
  ctx->release(d_stptr, sizeof(QpState *) * (size_t)states_cap);
  if (h_states) (void)hipHostFree(h_states);
  if (h_stptr) (void)hipHostFree(h_stptr);
  d_qps = nullptr;
  d_tiles = h_tiles = nullptr;
  d_ctrl = nullptr;
  d_flags = h_flags = nullptr;
  d_states = h_states = nullptr;
  d_stptr = h_stptr = nullptr;
  qps_cap = tiles_cap = flags_cap = states_cap = 0;
}

// Grows solve_batch's state staging buffers (called with the stream idle).
int BatchTables::ensure_states(int n_qps) {
  if (n_qps <= states_cap) return MPRGP_OK;
  const int cap = std::max(n_qps, 2 * states_cap);
  // the four buffers always have states_cap entries (or are all null), so a
  // failure part-way leaves nothing allocated and nothing counted
  auto drop = [&](int n) {
    ctx->release(d_states, 2 * sizeof(QpState) * (size_t)n);
    ctx->release(d_stptr, sizeof(QpState *) * (size_t)n);
    if (h_states) (void)hipHostFree(h_states);
    if (h_stptr) (void)hipHostFree(h_stptr);
    d_states = h_states = nullptr;
    d_stptr = h_stptr = nullptr;
    states_cap = 0;
  };
  drop(states_cap);
  int rc = ctx->alloc((void **)&d_states, 2 * sizeof(QpState) * (size_t)cap);
  if (!rc) rc = ctx->alloc((void **)&d_stptr, sizeof(QpState *) * (size_t)cap);
  hipError_t e = hipSuccess;
  if (!rc) e = hipHostMalloc((void **)&h_states, 2 * sizeof(QpState) * (size_t)cap, hipHostMallocDefault);
  if (!rc && e == hipSuccess) e = hipHostMalloc((void **)&h_stptr, sizeof(QpState *) * (size_t)cap, hipHostMallocDefault);
  if (!rc && e != hipSuccess) rc = fail(MPRGP_ENOMEM, "pinned staging of %d states: %s", cap, hipGetErrorString(e));
  if (rc) {
    drop(cap);
    return rc;
  }
  states_cap = cap;
  return MPRGP_OK;
}

// Grows the tables; graphs captured the old table pointers, so a regrowth
// drops them.
int BatchTables::ensure(int n_qps, int n_slots) {
  if (n_qps > flags_cap) {
    int cap = std::max(n_qps, 2 * flags_cap);
    ctx->release(d_flags, sizeof(int) * (size_t)flags_cap);
    if (h_flags) (void)hipHostFree(h_flags);
    d_flags = h_flags = nullptr;
    flags_cap = 0;
    if (int rc = ctx->alloc((void **)&d_flags, sizeof(int) * (size_t)cap)) return rc;
    hipError_t e = hipHostMalloc((void **)&h_flags, 2 * sizeof(int) * (size_t)cap, hipHostMallocDefault);
    if (e != hipSuccess) {
      h_flags = nullptr;
      ctx->release(d_flags, sizeof(int) * (size_t)cap);
      d_flags = nullptr;
      return fail(MPRGP_ENOMEM, "pinned flags of %d QPs: %s", cap, hipGetErrorString(e));
    }
    flags_cap = cap;
  }
  if (n_qps > qps_cap) {
    int cap = std::max(n_qps, 2 * qps_cap);
    ctx->release(d_qps, sizeof(QpDev) * (size_t)qps_cap);
    d_qps = nullptr;
    qps_cap = 0;
    if (int rc = ctx->alloc((void **)&d_qps, sizeof(QpDev) * (size_t)cap)) return rc;
    qps_cap = cap;
    drop_graphs();
  }
  if (n_slots > tiles_cap) {
    int cap = std::max(n_slots, 2 * tiles_cap);
    ctx->release(d_tiles, sizeof(Tile) * (size_t)tiles_cap);
    if (h_tiles) (void)hipHostFree(h_tiles);
    d_tiles = h_tiles = nullptr;
    tiles_cap = 0;
    if (int rc = ctx->alloc((void **)&d_tiles, sizeof(Tile) * (size_t)cap)) return rc;
    hipError_t e = hipHostMalloc((void **)&h_tiles, sizeof(Tile) * (size_t)cap, hipHostMallocDefault);
    if (e != hipSuccess) {
      h_tiles = nullptr;
      ctx->release(d_tiles, sizeof(Tile) * (size_t)cap);
      d_tiles = nullptr;
      return fail(MPRGP_ENOMEM, "pinned staging of %d tiles: %s", cap, hipGetErrorString(e));
    }
    tiles_cap = cap;
    drop_graphs();
  }
  return MPRGP_OK;
}

int BatchTables::graph(hipStream_t s, int ts, int steps, int grid, bool mixed, int variant, hipGraphExec_t *out,
                       int spw, bool c16) {
  auto key = std::make_tuple(ts, steps, grid, (int)mixed, variant, spw, (int)c16);
  auto it = graphs.find(key);
  if (it != graphs.end()) {
    *out = it->second;
    return MPRGP_OK;
  }
  // a new shape is captured with the stream idle (callers may have replays
  // in flight)
  MPRGP_HIP(hipStreamSynchronize(s));
  MPRGP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < steps; ++i) {
    hipError_t e = launch_step(ts, grid, s, d_qps, d_tiles, d_ctrl + 1, i & 1, mixed, variant, spw, c16);
    if (e != hipSuccess) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      return fail(MPRGP_EHIP, "capture failed: %s", hipGetErrorString(e));
    }
  }
  hipGraph_t g = nullptr;
  MPRGP_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(MPRGP_EHIP, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  graph_defs[key] = g;
  graphs[key] = ex;
  *out = ex;
  return MPRGP_OK;
}

// Wavefronts per workgroup: small batches use narrow workgroups so enough of
// them exist to fill the chip.  A QP's partial sums are grouped per tile, so
// its result is bit-identical across batches that use the same value.  From
// 512 slices on, 2 wavefronts: a lone QP's launch then re-reduces half as
// many tiles' partial sums (config 2, 696 slices: 14.0 against 15.0 us per
// power-method launch; a 224-slice QP: 9.7 against 8.9 us).
int pick_ts(const mprgp_params &p, long total_slices) {
  if (p.tiles == 1 || p.tiles == 2 || p.tiles == 4 || p.tiles == 8) return p.tiles;
  if (total_slices < 512) return 1;
  if (total_slices < 8192) return 2;
  return 4;
}

int tiles_for(int n_slices, int span) { return (n_slices + span - 1) / span; }

// Tiles of one QP (slot `slot` of the batch table): `span` = TS slices each,
// one per wavefront.
void append_tiles(int slot, const mprgp_qp *q, int span, std::vector<Tile> &tiles) {
  const int n_slices = q->A->n_slices;
  const int nt = tiles_for(n_slices, span);
  int local = 0;
  for (int s0 = 0; s0 < n_slices; s0 += span, ++local)
    tiles.push_back(Tile{q->st, q->part, slot, s0, std::min(span, n_slices - s0), local, nt, 0});
}

// Workgroups per launch: one per tile, rounded up to a power of two so that
// few distinct graphs get captured.
int round_grid(int n_tiles) {
  int g = 64;
  while (g < n_tiles) g *= 2;
  return g;
}

// The table a launch of `grid` workgroups reads, one slot per workgroup;
// unused slots hold qp = -1.  Workgroups go to the 8 XCDs of the MI355X
// round-robin (workgroup b to XCD b % 8), each XCD with its own L2: runs of
// kXcdRun consecutive tiles (neighbouring rows, which gather from each
// other) go to one XCD, and the runs rotate over the XCDs so that every XCD
// samples every QP.  On the 256-QP batch, runs of 4 measured 0.7% faster
// than single tiles, runs of 16 1% slower, and whole QPs per XCD 6% slower:
// per-QP work differs by op and one XCD's share then sets the launch time.
constexpr int kXcds = 8;
constexpr int kXcdRun = 4;
// layout 1 (per-QP chunks): each QP's tiles split into kXcds contiguous
// chunks, one per XCD, so an XCD gathers from one contiguous part of each
// QP's vectors; chunk sizes differ by at most one tile, the larger ones going
// to the XCDs with the fewest tiles so far.  Falls back to the runs when the
// chunks would not fit the table.
static bool layout_chunks(const std::vector<Tile> &tiles, int grid, Tile *table) {
  const int n = (int)tiles.size(), per = grid / kXcds;
  std::vector<int> len(kXcds, 0);
  std::vector<std::vector<int>> lists(kXcds);
  for (int i0 = 0; i0 < n;) {
    int i1 = i0;
    while (i1 < n && tiles[i1].qp == tiles[i0].qp) ++i1;
    const int nt = i1 - i0, base = nt / kXcds, rem = nt % kXcds;
    // XCDs in ascending order of their lists' length (ties by index)
    int order[kXcds];
    for (int x = 0; x < kXcds; ++x) order[x] = x;
    std::stable_sort(order, order + kXcds, [&](int a, int b) { return len[a] < len[b]; });
    int at = i0;
    for (int k = 0; k < kXcds; ++k) {
      const int x = order[k], sz = base + (k < rem ? 1 : 0);
      for (int j = 0; j < sz; ++j) lists[x].push_back(at++);
      len[x] += sz;
    }
    i0 = i1;
  }
  for (int x = 0; x < kXcds; ++x)
    if (len[x] > per) return false;
  for (int b = 0; b < grid; ++b) table[b] = Tile{nullptr, nullptr, -1, 0, 0, 0, 0, 0};
  for (int x = 0; x < kXcds; ++x)
    for (int pos = 0; pos < len[x]; ++pos) table[pos * kXcds + x] = tiles[lists[x][pos]];
  return true;
}

void layout_tiles(const std::vector<Tile> &tiles, int grid, Tile *table, int layout) {
  if (layout == 1 && layout_chunks(tiles, grid, table)) return;
  const int n = (int)tiles.size();
  // grid is a power of two >= 64, so c * kXcds divides it and every slot
  // below lands inside the table
  const int c = std::min(kXcdRun, grid / kXcds);
  for (int b = 0; b < grid; ++b) table[b] = Tile{nullptr, nullptr, -1, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const int xcd = (i / c) % kXcds, pos = (i / (c * kXcds)) * c + i % c;
    table[pos * kXcds + xcd] = tiles[i];
  }
}

int check_batch_qp(const mprgp_ctx *ctx, const mprgp_qp *q, const mprgp_params &p, int i) {
  if (!q) return fail(MPRGP_EBADINPUT, "qp %d is NULL", i);
  if (q->busy) return fail(MPRGP_ESTATE, "qp %d is busy in an engine", i);
  if (q->ctx != ctx) return fail(MPRGP_EBADINPUT, "qp %d belongs to another context", i);
  if (!q->has_rhs || !q->has_x) return fail(MPRGP_ESTATE, "qp %d: rhs and x must be set", i);
  if (!(p.tol == p.tol) || p.max_it < 0) return fail(MPRGP_EBADINPUT, "qp %d: bad tol/max_it", i);
  if (p.precision != MPRGP_FP64 && p.precision != MPRGP_MIXED)
    return fail(MPRGP_EBADINPUT, "qp %d: unknown precision %d", i, p.precision);
  if (p.variant != MPRGP_MONOTONIC && p.variant != MPRGP_CLASSIC)
    return fail(MPRGP_EBADINPUT, "qp %d: unknown variant %d", i, p.variant);
  if (p.precision == MPRGP_MIXED && p.variant != MPRGP_MONOTONIC)
    return fail(MPRGP_EBADINPUT, "qp %d: the mixed form runs the monotone variant", i);
  if (!(p.slices_per_wave == 0 || p.slices_per_wave == 1 || p.slices_per_wave == 2))
    return fail(MPRGP_EBADINPUT, "qp %d: slices_per_wave must be 1 or 2 (0: 1)", i);
  if (p.from_x0 != 0 && p.from_x0 != 1) return fail(MPRGP_EBADINPUT, "qp %d: from_x0 must be 0 or 1", i);
  return MPRGP_OK;
}

// Launch bound of one solve: every op sequence of an iteration is at most 8
// launches, plus the power method and the refinement sweeps.
long solve_step_bound(const mprgp_params &p) {
  const bool mixed = p.precision == MPRGP_MIXED;
  const bool power = !(p.alpha_bar > 0.0);
  // mixed: the sweeps, their final R_START, and a possible FP64 finish
  const long sweeps = mixed ? (p.max_refine > 0 ? p.max_refine : 8) + 2 : 1;
  return sweeps * (8L * ((long)p.max_it + 4) + 8) + (power ? 2L * 1002 + 4 : 0) + 16;
}

// Bytes of device memory a solve of `q` works on (matrix, its FP32 copy for
// the mixed form, the QP's vectors).
size_t solve_bytes(const mprgp_qp *q, const mprgp_params &p) {
  size_t b = q->A->bytes + q->bytes;
  if (p.precision == MPRGP_MIXED) {
    b += sizeof(float) * 9 * (size_t)q->A->slots;
    b += (7 * 3 * (size_t)q->n + (size_t)q->n) * sizeof(double) + (3 * 3 * (size_t)q->n + 64) * sizeof(float);
  }
  return b;
}

// The device descriptor and the initial state of a QP entering a batch (host
// structs only; the mixed form's FP32 storage must exist).
void describe_qp(const mprgp_qp *q, const mprgp_params &p, int n_tiles, int *done_flag, int tag, QpDev &d,
                 QpState &st) {
  std::memset(&d, 0, sizeof(d));
  d.n_nodes = q->n;
  d.n_slices = q->A->n_slices;
  d.done_flag = done_flag;
  d.tag = tag;
  d.sl_off = q->A->d_sl_off;
  d.col = q->A->d_col;
  d.col16 = q->A->d_col16;
  d.d.val = q->A->d_val;
  d.d.dinv = q->A->d_dinv;
  d.d.x = q->x; d.d.b = q->b; d.d.g = q->g; d.d.p = q->p; d.d.z = q->z; d.d.phi = q->phi;
  d.d.beta = q->beta; d.d.y = q->y; d.d.ap = q->ap; d.d.py = q->py;
  d.d.p2 = q->p2; d.d.beta2 = q->beta2;
  d.d.cn = q->cn; d.d.cc = q->cc; d.d.cj = q->cj; d.d.cs = q->cs;
  d.ev = q->ev; d.pv = q->pv; d.pw = q->pw; d.pu = q->pu; d.pz = q->pz;
  d.cf = q->cf; d.face = q->face;
  const bool mixed = p.precision == MPRGP_MIXED;
  if (mixed) {
    d.f.val = q->A->d_val32;
    d.f.dinv = q->A->d_dinv;
    d.f.x = q->fx; d.f.b = q->fb; d.f.g = q->fg; d.f.p = q->fp; d.f.z = q->fz; d.f.phi = q->fphi;
    d.f.beta = q->fbeta; d.f.y = q->fy; d.f.ap = q->fap; d.f.py = q->fpy; d.f.p2 = q->fp2;
    d.f.beta2 = q->fbeta2;
    d.f.cn = q->cn; d.f.cj = q->cj; d.f.cs = q->cs; d.f.cc = q->fcc;
  }
  std::memset(&st, 0, sizeof(st));
  st.n_tiles = n_tiles;
  st.max_it = p.max_it;
  st.tol = p.tol < 1e-30 ? 1e-30 : p.tol;
  st.gamma = 1.0;
  st.code = -1;
  st.has_rows = q->m > 0;
  st.precond = p.preconditioned ? 1 : 0;
  st.variant = p.variant;
  const bool power = !(p.alpha_bar > 0.0);
  st.alpha_bar = power ? 0.0 : p.alpha_bar;
  st.lambda_max = power ? 0.0 : 2.0 / p.alpha_bar;
  st.power_eps = p.power_tol > 0.0 ? p.power_tol : p.tol;
  st.pow_next = power ? OP_POW_IT : (mixed ? OP_R_START : OP_INIT);
  st.op = p.project_first ? OP_PROJX : st.pow_next;
  st.fresh = 1;
  if (mixed) {
    st.mixed = 1;
    st.tol_outer = st.tol;
    st.max_refine = p.max_refine > 0 ? p.max_refine : 8;
    st.inner_rel = p.inner_tol > 0.0 ? p.inner_tol : 1e-3;
  }
}

// describe_qp, and the QP's set-up enqueued on stream `s` (FP32 copies, power
// start vector, tangent-Jacobi factors).  The caller uploads `st` into both
// state copies.
int prepare_qp(mprgp_qp *q, const mprgp_params &p, hipStream_t s, int n_tiles, int *done_flag, int tag,
               QpDev &d, QpState &st, bool rows) {
  if (q->ctx->stream != s) {
    // work the QP's own context queued (reset_x, a collider, an assembly
    // not yet synchronised) precedes the solve on stream s
    std::lock_guard<std::recursive_mutex> lk(q->ctx->mu);
    MPRGP_HIP(hipEventRecord(q->ctx->ready, q->ctx->stream));
    MPRGP_HIP(hipStreamWaitEvent(s, q->ctx->ready, 0));
  }
  if (p.precision == MPRGP_MIXED) {
    if (int rc = ensure_matrix32(q->A, s)) return rc;
    if (int rc = ensure_qp32(q, s)) return rc;
  }
  describe_qp(q, p, n_tiles, done_flag, tag, d, st);
  if (p.from_x0) {  // the warm start the last set_x stored, as mprgp_qp_reset_x
    if (rows) MPRGP_HIP(hipMemcpyAsync(q->x, q->x0, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, s));
    else d.x0 = q->x0;
  }
  if (!(p.alpha_bar > 0.0)) {
    if (p.use_ev && q->ev_valid) {
      MPRGP_HIP(hipMemcpyAsync(q->pv, q->ev, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, s));
    } else {
      // the deterministic start vector in user node order, normalised on the
      // device, then in internal order
      const size_t n3 = 3 * (size_t)q->n;
      MPRGP_HIP(launch_power_start(s, q->stage, n3, power_start_norm(n3)));
      MPRGP_HIP(launch_gather_nodes(s, q->stage, q->pv, q->A->d_perm, q->n));
    }
  }
  // tangent-Jacobi factors depend on the current matrix values
  if (rows) MPRGP_HIP(launch_setup_rows(s, q->A->d_dinv, q->cn, q->cf, q->cs, q->n));
  return MPRGP_OK;
}

// The state copy that reached DONE (both do one launch later).
const QpState &final_state(const QpState both[2]) {
  return (both[1].op == OP_DONE && (both[0].op != OP_DONE || both[1].steps > both[0].steps)) ? both[1]
                                                                                              : both[0];
}

int fill_stats(mprgp_qp *q, const QpState both[2], float ms, long launched, mprgp_stats *out) {
  const QpState &st = final_state(both);
  if (st.ev_valid) q->ev_valid = 1;
  std::memcpy(q->opcnt, st.opcnt, sizeof(st.opcnt));
  if (out) {
    mprgp_stats &o = *out;
    std::memset(&o, 0, sizeof(o));
    o.code = st.code;
    o.refinements = (st.mixed || st.fp64_finish) ? st.refine : 0;
    o.fp64_finish = st.fp64_finish;
    o.iterations = st.iteration;
    o.num_cg = st.num_cg;
    o.num_exp = st.num_exp;
    o.num_prop = st.num_prop;
    o.active = st.active;
    o.power_iters = st.power_it;
    o.ev_valid = q->ev_valid;
    o.steps = st.steps;
    o.residual = st.residual;
    o.alpha_bar = st.alpha_bar;
    o.lambda_max = st.lambda_max;
    o.elapsed_ms = ms;
    o.launches = (int)launched;  // step-kernel launches issued for the whole batch
    const double lo = q->A->lambda_min_lower;
    o.error_bound = (lo > 0.0 && std::isfinite(st.residual)) ? st.residual / lo : -1.0;
  }
  if (!std::isfinite(st.residual)) return MPRGP_ENAN;
  if (st.op != OP_DONE) return MPRGP_EHIP;
  return st.code == 0 ? MPRGP_OK : MPRGP_NOT_CONVERGED;
}

}  // namespace mprgp

static int run_batch(mprgp_ctx *ctx, int count, mprgp_qp *const *qps, const mprgp_params *params,
                     int stride, mprgp_stats *stats) {
  if (!ctx || count <= 0 || !qps || !params) return fail(MPRGP_EBADINPUT, "bad batch arguments");
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  for (int i = 0; i < count; ++i)
    if (int rc = check_batch_qp(ctx, qps[i], params[stride ? i : 0], i)) return rc;
  {  // a QP may appear once: sort the handles and compare neighbours
    std::vector<std::pair<const mprgp_qp *, int>> h(count);
    for (int i = 0; i < count; ++i) h[i] = {qps[i], i};
    std::sort(h.begin(), h.end());
    for (int i = 1; i < count; ++i)
      if (h[i].first == h[i - 1].first)
        return fail(MPRGP_EBADINPUT, "qp %d appears twice in the batch", std::max(h[i].second, h[i - 1].second));
  }
  bool any_mixed = false, all_c16 = true;
  long total_slices = 0, max_steps = 0;
  std::vector<int> n_slices(count);
  for (int i = 0; i < count; ++i) {
    const mprgp_params &p = params[stride ? i : 0];
    any_mixed |= p.precision == MPRGP_MIXED;
    all_c16 &= qps[i]->A->d_col16 != nullptr;
    total_slices += (n_slices[i] = qps[i]->A->n_slices);
    max_steps = std::max(max_steps, solve_step_bound(p));
  }
  const int ts = pick_ts(params[0], total_slices);
  const int spw = slices_per_wave_of(params[0]);
  const int span = ts * spw;
  const bool c16 = use_col16(all_c16, spw);
  std::vector<Tile> tiles;
  for (int i = 0; i < count; ++i) append_tiles(i, qps[i], span, tiles);
  BatchTables &tab = ctx->tab;
  const int n_tiles = (int)tiles.size();
  int grid = round_grid(n_tiles);
  if (int rc = tab.ensure(count, grid)) return rc;
  if (int rc = tab.ensure_states(count)) return rc;
  const int layout = params[0].xcd_layout;
  layout_tiles(tiles, grid, tab.h_tiles, layout);
  hipStream_t s = ctx->stream;
  std::vector<QpDev> devs(count);
  QpState *states = tab.h_states;  // pinned; [i][2]
  int max_nodes = 0;
  for (int i = 0; i < count; ++i) {
    QpState &st = states[2 * (size_t)i];
    if (int rc = prepare_qp(qps[i], params[stride ? i : 0], s, tiles_for(n_slices[i], span), tab.d_flags + i, 1,
                            devs[i], st, count == 1 || count > 65535))
      return rc;
    max_nodes = std::max(max_nodes, qps[i]->n);
    states[2 * (size_t)i + 1] = st;
    tab.h_stptr[i] = qps[i]->st;
  }
  // every QP's two state copies in one upload and one scatter (a lone QP:
  // one direct copy)
  if (count == 1) {
    MPRGP_HIP(hipMemcpyAsync(qps[0]->st, states, 2 * sizeof(QpState), hipMemcpyHostToDevice, s));
  } else {
    MPRGP_HIP(hipMemcpyAsync(tab.d_states, states, 2 * sizeof(QpState) * (size_t)count, hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(tab.d_stptr, tab.h_stptr, sizeof(QpState *) * (size_t)count, hipMemcpyHostToDevice, s));
    MPRGP_HIP(launch_scatter_states(s, tab.d_states, tab.d_stptr, count));
  }
  const int running0 = count;
  MPRGP_HIP(hipMemcpyAsync(tab.d_qps, devs.data(), sizeof(QpDev) * count, hipMemcpyHostToDevice, s));
  // the rows' tangent-Jacobi factors of every QP (they follow the current
  // matrix values) in one launch
  if (count > 1 && count <= 65535) MPRGP_HIP(launch_setup_rows_batch(s, tab.d_qps, count, max_nodes));
  MPRGP_HIP(hipMemcpyAsync(tab.d_tiles, tab.h_tiles, sizeof(Tile) * grid, hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemcpyAsync(tab.d_ctrl + 1, &running0, sizeof(int), hipMemcpyHostToDevice, s));
  MPRGP_HIP(hipMemsetAsync(tab.d_flags, 0, sizeof(int) * count, s));
  MPRGP_HIP(hipStreamSynchronize(s));

  int steps = params[0].graph_steps > 0 ? params[0].graph_steps : 16;
  steps += steps & 1;  // even, so every replay starts at parity 0
  // graph_steps < 0 (or MPRGP_NO_GRAPH set): plain stream launches, e.g. for
  // profilers that do not follow graph replays
  const char *ng = getenv("MPRGP_NO_GRAPH");
  const bool use_graph = params[0].graph_steps >= 0 && !(ng && ng[0] && ng[0] != '0');
  if (params[0].graph_steps < 0) steps = -params[0].graph_steps + (params[0].graph_steps & 1);
  // the kernel variant for the table's tiles and the bytes of their matrices
  // (same results; few tiles: the tail variant, large matrices: streamed)
  auto live_bytes = [&](const std::vector<char> *keep) {
    double b = 0.0;
    for (int i = 0; i < count; ++i)
      if (!keep || (*keep)[i]) b += matrix_stream_bytes(qps[i]->A, params[stride ? i : 0].precision == MPRGP_MIXED);
    return b;
  };
  const int tail_tiles = tail_tiles_of(params[0]);
  int variant = step_variant(n_tiles, live_bytes(nullptr), tail_tiles);
  hipGraphExec_t ex = nullptr;
  if (use_graph)
    if (int rc = tab.graph(s, ts, steps, grid, any_mixed, variant, &ex, spw, c16)) return rc;
  int in_table = count;  // QPs whose tiles are in the table

  MPRGP_HIP(hipEventRecord(ctx->t0, s));
  long launched = 0;
  int k = 0;
  bool done = false;
  auto enqueue = [&](int slot) -> int {
    if (use_graph) {
      MPRGP_HIP(hipGraphLaunch(ex, s));
    } else {
      for (int i = 0; i < steps; ++i)
        MPRGP_HIP(launch_step(ts, grid, s, tab.d_qps, tab.d_tiles, tab.d_ctrl + 1, i & 1, any_mixed, variant, spw,
                              c16));
    }
    launched += steps;
    MPRGP_HIP(hipMemcpyAsync(ctx->h_running + slot, tab.d_ctrl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    if (count > 1)
      MPRGP_HIP(hipMemcpyAsync(tab.h_flags + (size_t)slot * tab.flags_cap, tab.d_flags, sizeof(int) * count,
                               hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipEventRecord(ctx->q[slot], s));
    return MPRGP_OK;
  };
  if (int rc = enqueue(0)) return rc;
  std::vector<char> keep(count, 1);
  long shrunk_at = -2;  // the poll of the last table shrink
  while (!done) {
    // two replays in flight: the next one is queued before this poll's wait
    if (int rc = enqueue((k + 1) & 1)) return rc;
    MPRGP_HIP(hipEventSynchronize(ctx->q[k & 1]));
    const int running = ctx->h_running[k & 1];
    const int *flags = tab.h_flags + (size_t)(k & 1) * tab.flags_cap;
    if (running == 0) done = true;
    // Drop the tiles of finished QPs once a quarter of the table is idle,
    // without draining the stream: the flags of this poll name the QPs done
    // by this replay (those the replay in flight finishes stay one replay
    // longer and idle), and the new table is uploaded behind the replay in
    // flight, which still reads the old one.  The staging buffer is written
    // again only from the second poll on, whose replay follows the last
    // upload.
    if (!done && count > 1 && 4 * running < 3 * in_table && k >= shrunk_at + 2) {
      int live = 0;
      for (int i = 0; i < count; ++i) {
        keep[i] = flags[i] == 0;
        live += keep[i];
      }
      if (live > 0 && live < in_table) {
        std::vector<Tile> kept;
        for (int i = 0; i < count; ++i)
          if (keep[i]) append_tiles(i, qps[i], span, kept);
        grid = round_grid((int)kept.size());  // <= the table's capacity: the tile count only shrinks
        layout_tiles(kept, grid, tab.h_tiles, layout);
        MPRGP_HIP(hipMemcpyAsync(tab.d_tiles, tab.h_tiles, sizeof(Tile) * grid, hipMemcpyHostToDevice, s));
        in_table = live;
        shrunk_at = k;
        variant = step_variant((int)kept.size(), live_bytes(&keep), tail_tiles);
        if (use_graph)
          if (int rc = tab.graph(s, ts, steps, grid, any_mixed, variant, &ex, spw, c16)) return rc;
      }
    }
    ++k;
    if (!done && launched > max_steps + 2L * steps) {
      (void)hipStreamSynchronize(s);
      return fail(MPRGP_EHIP, "solve did not finish within %ld device steps", launched);
    }
  }
  MPRGP_HIP(hipEventRecord(ctx->t1, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  MPRGP_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));

  // every QP's two state copies in one gather and one readback
  if (count == 1) {
    MPRGP_HIP(hipMemcpyAsync(tab.h_states, qps[0]->st, 2 * sizeof(QpState), hipMemcpyDeviceToHost, s));
  } else {
    MPRGP_HIP(launch_gather_states(s, tab.d_stptr, tab.d_states, count));
    MPRGP_HIP(hipMemcpyAsync(tab.h_states, tab.d_states, 2 * sizeof(QpState) * (size_t)count, hipMemcpyDeviceToHost, s));
  }
  MPRGP_HIP(hipStreamSynchronize(s));
  int worst = MPRGP_OK;
  for (int i = 0; i < count; ++i) {
    const QpState *both = tab.h_states + 2 * (size_t)i;
    int rc = fill_stats(qps[i], both, ms, launched, stats ? stats + i : nullptr);
    if (rc == MPRGP_ENAN || rc == MPRGP_EHIP) worst = rc;
    else if (rc == MPRGP_NOT_CONVERGED && worst == MPRGP_OK) worst = rc;
  }
  return worst;
}

extern "C" int mprgp_qp_solve(mprgp_qp *q, const mprgp_params *params, mprgp_stats *stats) {
  MPRGP_TRY
  if (!q) return fail(MPRGP_EBADINPUT, "NULL qp");
  mprgp_params def;
  if (!params) {
    mprgp_params_default(&def);
    params = &def;
  }
  return run_batch(q->ctx, 1, &q, params, 0, stats);
  MPRGP_CATCH
}

extern "C" int mprgp_solve_batch(mprgp_ctx *ctx, int count, mprgp_qp *const *qps, const mprgp_params *params,
                                 int params_stride, mprgp_stats *stats) {
  MPRGP_TRY
  mprgp_params def;
  if (!params) {
    mprgp_params_default(&def);
    params = &def;
    params_stride = 0;
  }
  return run_batch(ctx, count, qps, params, params_stride ? 1 : 0, stats);
  MPRGP_CATCH
}

extern "C" int mprgp_copy_probe(mprgp_ctx *ctx, size_t bytes, int iters, double *gbps) {
  MPRGP_TRY
  if (!ctx || !gbps || iters <= 0 || bytes < 4096) return fail(MPRGP_EBADINPUT, "bad probe arguments");
  bytes &= ~(size_t)15;
  MPRGP_HIP(hipSetDevice(ctx->device));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  void *a = nullptr, *b = nullptr;
  if (int rc = ctx->alloc(&a, bytes)) return rc;
  if (int rc = ctx->alloc(&b, bytes)) {
    ctx->release(a, bytes);
    return rc;
  }
  hipStream_t s = ctx->stream;
  int rc = MPRGP_OK;
  hipError_t e = hipMemsetAsync(a, 0, bytes, s);
  if (!e) e = launch_copy_stream(s, a, b, bytes);  // warm-up
  if (!e) e = hipEventRecord(ctx->t0, s);
  for (int i = 0; i < iters && !e; ++i) e = launch_copy_stream(s, (i & 1) ? b : a, (i & 1) ? a : b, bytes);
  if (!e) e = hipEventRecord(ctx->t1, s);
  if (!e) e = hipStreamSynchronize(s);
  float ms = 0.f;
  if (!e) e = hipEventElapsedTime(&ms, ctx->t0, ctx->t1);
  if (e) rc = fail(MPRGP_EHIP, "copy probe: %s", hipGetErrorString(e));
  else *gbps = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  ctx->release(a, bytes);
  ctx->release(b, bytes);
  return rc;
  MPRGP_CATCH
}

// The one-shot drop-in on a context of its own; J by rows (CSR) or by
// columns (CSC, Eigen's column-major storage).
int mprgp::decoupled_solve(int device, int n, const int *a_outer, const int *a_inner, const double *a_val,
                           const double *B, int m, int j_col_major, const int *j_outer, const int *j_inner,
                           const double *j_val, const double *c, double *x, double tol, int max_it, double *ev,
                           int *ev_valid, mprgp_stats *stats) {
  if (!B || !x) return fail(MPRGP_EBADINPUT, "NULL B or x");
  mprgp_ctx *ctx = nullptr;
  int rc = mprgp_ctx_create(device, 0, &ctx);
  if (rc) return rc;
  mprgp_matrix *A = nullptr;
  mprgp_qp *q = nullptr;
  rc = mprgp_matrix_create_csr(ctx, n, a_outer, a_inner, a_val, &A);
  if (!rc) rc = mprgp_qp_create(A, &q);
  if (!rc) rc = mprgp_qp_set_rows_sparse(q, m, j_col_major, j_outer, j_inner, j_val, c);
  if (!rc) rc = mprgp_qp_set_rhs(q, B);
  if (!rc) rc = mprgp_qp_set_x(q, x);
  if (!rc && ev && ev_valid && *ev_valid) rc = mprgp_qp_set_ev(q, ev);
  int code = rc;
  if (!rc) {
    mprgp_params p;
    mprgp_params_default(&p);
    p.tol = tol;
    p.max_it = max_it;
    p.project_first = 1;
    code = mprgp_qp_solve(q, &p, stats);
    if (code == MPRGP_OK || code == MPRGP_NOT_CONVERGED) {
      int rc2 = mprgp_qp_get_x(q, x);
      if (!rc2 && ev && ev_valid) rc2 = mprgp_qp_get_ev(q, ev, ev_valid);
      if (rc2) code = rc2;
    }
  }
  mprgp_qp_destroy(q);
  mprgp_matrix_destroy(A);
  mprgp_ctx_destroy(ctx);
  return code;
}

extern "C" int mprgp_decoupled_solve(int device, int n, const int *a_outer, const int *a_inner,
                                     const double *a_val, const double *B, int m, const int *j_outer,
                                     const int *j_inner, const double *j_val, const double *c, double *x,
                                     double tol, int max_it, double *ev, int *ev_valid, mprgp_stats *stats) {
  MPRGP_TRY
  return mprgp::decoupled_solve(device, n, a_outer, a_inner, a_val, B, m, /*j_col_major*/ 1, j_outer, j_inner,
                                j_val, c, x, tol, max_it, ev, ev_valid, stats);
  MPRGP_CATCH
}

// MPRGPDecoupledCon::solve(file_name, x, ...) (MPRGPSolver.h:576-589): the
// file's A, B, J, c and x0, then the one-shot drop-in from x0.
extern "C" int mprgp_decoupled_solve_file(int device, const char *file_name, double *x, int n_x, int *n_out,
                                          double tol, int max_it, double *ev, int *ev_valid,
                                          mprgp_stats *stats) {
  MPRGP_TRY
  QpCsr q;
  if (int rc = read_qp_csr(file_name, q)) return rc;
  if (n_out) *n_out = q.n;
  if (!x || n_x < q.n) return fail(MPRGP_EBADINPUT, "x holds %d values, the QP has %d", n_x, q.n);
  std::copy(q.x0.begin(), q.x0.end(), x);
  return mprgp::decoupled_solve(device, q.n, q.ap.data(), q.ac.data(), q.av.data(), q.B.data(), q.m,
                                /*j_col_major*/ 0, q.jp.data(), q.jc.data(), q.jv.data(), q.c.data(), x, tol,
                                max_it, ev, ev_valid, stats);
  MPRGP_CATCH
}
