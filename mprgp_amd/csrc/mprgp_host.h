// mprgp_host.h -- host-side objects behind the C ABI (internal header).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mprgp_amd.h"
#include "mprgp_dev.h"
#include "mprgp_error.h"

struct mprgp_ctx;
struct mprgp_matrix;

namespace mprgp {

// the one-shot drop-in behind mprgp_decoupled_solve and the file overload;
// J by rows (j_col_major = 0) or by columns (1)
int decoupled_solve(int device, int n, const int *a_outer, const int *a_inner, const double *a_val,
                    const double *B, int m, int j_col_major, const int *j_outer, const int *j_inner,
                    const double *j_val, const double *c, double *x, double tol, int max_it, double *ev,
                    int *ev_valid, mprgp_stats *stats);

#define MPRGP_HIP(call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::mprgp::fail(MPRGP_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                           \
  } while (0)

// The power method's deterministic start vector, replacing the reference's
// unseeded setRandom() (MPRGPSolver.h:47-50): element i is 2u - 1 for u the
// top 53 bits of the i-th splitmix64 output from seed 0x5EED, i.e. of the
// state 0x5EED + (i + 1) * 0x9E3779B97F4A7C15 mixed; the vector is then
// divided by its norm (power_start_norm, mprgp_capi.cpp).  Every operation
// is exact except the final division, which rounds the same on host and
// device.
__host__ __device__ inline double power_start_raw(uint64_t i) {
  uint64_t z = 0x5EEDULL + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}
double power_start_norm(size_t n);

// Rows J x >= c given over the user's DoFs (CSR, m rows): J's columns in a
// QP's internal DoF order (jc, same entry order), and J^T as a CSR of
// 3 * nodes rows whose entries run over J's rows in ascending order.
// Validates the arrays (EBADINPUT on empty rows, bad columns, non-finite
// values).
struct RowsInternal {
  std::vector<int> jc;
  std::vector<int> tptr, tc;
  std::vector<double> tv;
};
int rows_to_internal(const mprgp_matrix *A, int m, const int *j_ptr, const int *j_col, const double *j_val,
                     const double *c, RowsInternal &out);

// kernel launch wrappers (mprgp_kernels.hip)
// coupled-row solve: out = M x for a CSR M of `rows` rows; out = a x + b y;
// out = x y (elementwise)
hipError_t launch_csr_mul(hipStream_t s, int rows, const int *ptr, const int *col, const double *val,
                          const double *x, double *out);
hipError_t launch_axpby(hipStream_t s, size_t n, double a, const double *x, double b, const double *y, double *out);
// k <= kMaxCsrVecs products out[j] = C x[j] with one CSR matrix C in one
// launch, each the bits launch_csr_mul gives
constexpr int kMaxCsrVecs = 4;
hipError_t launch_csr_mul_n(hipStream_t s, int rows, const int *ptr, const int *col, const double *val, int k,
                            const double *const *x, double *const *out);
hipError_t launch_vmul(hipStream_t s, size_t n, const double *x, const double *y, double *out);
// out[i] = power_start_raw(i) / nr for i < n (out as is when nr = 0)
hipError_t launch_power_start(hipStream_t s, double *out, size_t n, double nr);
// variant: a StepVariant (mprgp_dev.h), from step_variant() below
// spw: slices per wavefront (1 or 2)
// c16: every QP of the table has 16-bit column indices (ignored with spw 2)
hipError_t launch_step(int ts, int grid, hipStream_t s, const QpDev *qps, const Tile *tiles,
                       int *running, int parity, bool mixed, int variant = kStepCached, int spw = 1,
                       bool c16 = false);
hipError_t launch_gather_nodes(hipStream_t s, const double *in, double *out, const int *perm, int n);
// setup_rows for `count` QPs of a device QpDev table (count <= 65535)
hipError_t launch_setup_rows_batch(hipStream_t s, const QpDev *qps, int count, int max_nodes);
// the two state copies of n QPs: staging [i][2] <-> each QP's own (ptr[i])
hipError_t launch_scatter_states(hipStream_t s, const QpState *src, QpState *const *dst, int n);
hipError_t launch_gather_states(hipStream_t s, QpState *const *src, QpState *dst, int n);
hipError_t launch_scatter_nodes(hipStream_t s, const double *in, double *out, const int *perm, int n);
hipError_t launch_setup_diag(hipStream_t s, const int *sl_off, const int *col, const double *val,
                             double *dinv, int n);
hipError_t launch_setup_rows(hipStream_t s, const double *dinv, const double *cn,
                             const unsigned char *cf, double *cs, int n);
hipError_t launch_spmv_plain(hipStream_t s, const int *sl_off, const int *col, const double *val,
                             const double *x, double *y, int n, int accumulate);
hipError_t launch_collide_planes(hipStream_t s, const double *x, const double *planes, int n_planes,
                                 double *cn, double *cc, double *cj, unsigned char *cf, int *count, int n);
hipError_t launch_get_diag(hipStream_t s, const int *sl_off, const int *col, const double *val, double *diag,
                           int n);
// a.b into out[0] (device); part holds one double per 256 entries
hipError_t launch_dot(hipStream_t s, const double *a, const double *b, size_t n, double *part, double *out);
// k <= kMaxDots dot products a[j].b[j] into out[0..k) with one kernel pair;
// each has the bits launch_dot gives it.  part holds k doubles per 256 entries
constexpr int kMaxDots = 8;
hipError_t launch_dots(hipStream_t s, int k, const double *const *a, const double *const *b, size_t n,
                       double *part, double *out);
// part (if non-null) must hold one double per 256 nodes; result receives the reduction
hipError_t launch_projector(hipStream_t s, const ProjArgs &a, double *result);
hipError_t launch_to_f32(hipStream_t s, const double *in, float *out, size_t n);
hipError_t launch_to_u16(hipStream_t s, const int *in, unsigned short *out, size_t n);
hipError_t launch_copy_stream(hipStream_t s, const void *in, void *out, size_t bytes);

// Builds the SELL-64 device matrix from a node-order block CSR (blocks may be
// NULL for a zero matrix of that pattern).
int build_matrix(mprgp_ctx *ctx, int n_nodes, const int *rowptr, const int *colidx,
                 const double *blocks, mprgp_matrix **out);

// Device launch tables of one batch driver (a context's solve_batch, or an
// engine): QP descriptors, tiles, control words, done flags, captured graphs.
struct BatchTables {
  mprgp_ctx *ctx = nullptr;
  QpDev *d_qps = nullptr;
  int qps_cap = 0;
  Tile *d_tiles = nullptr;  // one slot per workgroup of a launch (layout_tiles)
  Tile *h_tiles = nullptr;  // pinned staging of a table upload, tiles_cap slots
  int tiles_cap = 0;
  int *d_ctrl = nullptr;   // [1] = running QPs
  int *d_flags = nullptr;  // per batch slot: tag of the QP that reached DONE
  int *h_flags = nullptr;  // pinned, [2][flags_cap]: d_flags read back with each poll
  int flags_cap = 0;
  // solve_batch's QP states in one staging copy each way: [qp][2] states,
  // device and pinned host, and each QP's own state address
  QpState *d_states = nullptr, *h_states = nullptr;
  QpState **d_stptr = nullptr, **h_stptr = nullptr;
  int states_cap = 0;
  int ensure_states(int n_qps);
  // key: wavefronts per workgroup, launches, grid, mixed, kernel variant,
  // slices per wavefront, 16-bit columns
  std::map<std::tuple<int, int, int, int, int, int, int>, hipGraphExec_t> graphs;
  std::map<std::tuple<int, int, int, int, int, int, int>, hipGraph_t> graph_defs;
  int init(mprgp_ctx *c);
  int ensure(int n_qps, int n_slots);
  int graph(hipStream_t s, int ts, int steps, int grid, bool mixed, int variant, hipGraphExec_t *out, int spw = 1,
            bool c16 = false);
  void drop_graphs();
  void destroy();
};

// Pieces of the batch solve shared by solve_batch and the engine.
int check_batch_qp(const mprgp_ctx *ctx, const mprgp_qp *q, const mprgp_params &p, int i);
int pick_ts(const mprgp_params &p, long total_slices);
// slices each wavefront streams per launch (params.slices_per_wave: 1 or 2)
inline int slices_per_wave_of(const mprgp_params &p) { return p.slices_per_wave == 2 ? 2 : 1; }
// Whether a launch table's SpMVs read 16-bit column indices: every QP's
// matrix has them (all_have) and one slice per wavefront.  MPRGP_INDEX32=1
// forces the 32-bit ones (the two give the same results, bit for bit).
inline bool use_col16(bool all_have, int spw) {
  const char *e = getenv("MPRGP_INDEX32");
  return all_have && spw == 1 && !(e && e[0] && e[0] != '0');
}
void append_tiles(int slot, const mprgp_qp *q, int span, std::vector<Tile> &tiles);
int tiles_for(int n_slices, int span);
int round_grid(int n_tiles);
// Launches with at most this many tiles (one workgroup per CU of an MI355X)
// use the step kernel's tail variant, unless the call's tail_tiles says
// otherwise.
constexpr int kTailTiles = 256;
inline int tail_tiles_of(const mprgp_params &p) { return p.tail_tiles > 0 ? p.tail_tiles : kTailTiles; }
// Launches whose matrices (SELL slots of the live QPs, at each QP's storage
// precision) exceed the MI355X's 256 MB last-level cache stream them
// nontemporal.  MPRGP_STREAM_BYTES, when set, replaces the threshold (a
// tuning knob; the variants give the same results).
constexpr double kStreamBytes = 256.0 * 1024 * 1024;
inline double stream_threshold() {
  const char *e = getenv("MPRGP_STREAM_BYTES");
  return (e && e[0]) ? atof(e) : kStreamBytes;
}
inline int step_variant(int n_tiles, double matrix_bytes, int tail_tiles = kTailTiles) {
  if (n_tiles <= tail_tiles) return kStepTail;
  return matrix_bytes > stream_threshold() ? kStepStream : kStepCached;
}
// layout 0: runs of kXcdRun tiles rotating over the XCDs; 1: per-QP chunks
// (one contiguous chunk of every QP per XCD).  Results do not depend on it.
void layout_tiles(const std::vector<Tile> &tiles, int grid, Tile *table, int layout = 0);
long solve_step_bound(const mprgp_params &p);
size_t solve_bytes(const mprgp_qp *q, const mprgp_params &p);
// the descriptor and initial state only (no device work; a mixed-form QP's
// FP32 storage must exist)
void describe_qp(const mprgp_qp *q, const mprgp_params &p, int n_tiles, int *done_flag, int tag, QpDev &d,
                 QpState &st);
// rows = false: the caller runs the rows' tangent-Jacobi set-up and the
// from_x0 restart itself (solve_batch: one launch for the whole batch, which
// copies the x0 that d.x0 names).
int prepare_qp(mprgp_qp *q, const mprgp_params &p, hipStream_t s, int n_tiles, int *done_flag, int tag,
               QpDev &d, QpState &st, bool rows = true);
int fill_stats(mprgp_qp *q, const QpState both[2], float ms, long launched, mprgp_stats *out);
// The no-alloc solve's per-device scratch address (hipGetSymbolAddress with
// the device current), looked up once per device at context creation.
int noalloc_setup(int device);

}  // namespace mprgp

struct mprgp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  size_t in_use = 0;
  std::mutex amu;            // guards budget accounting (alloc/release)
  std::recursive_mutex mu;   // serialises calls that use `stream`
  mprgp::BatchTables tab;    // solve_batch's launch tables
  int *h_running = nullptr;  // pinned, 2 slots
  void *h_final = nullptr;   // pinned: a QP's two state copies, read back by the no-alloc solve
  hipEvent_t t0 = nullptr, t1 = nullptr, q[2] = {nullptr, nullptr};
  hipEvent_t ready = nullptr;  // marks the stream's work for solves on another stream
  // scratch shared by calls on this context (serialised by `mu`)
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  int scratch_reserve(size_t bytes);

  int alloc(void **p, size_t bytes);
  void release(void *p, size_t bytes);
};

struct mprgp_matrix {
  mprgp_ctx *ctx = nullptr;
  int n_nodes = 0;
  int n_slices = 0;
  long long nb = 0;      // real blocks
  long long slots = 0;   // SELL slots including padding
  std::vector<int> perm;   // internal -> user node
  std::vector<int> iperm;  // user -> internal node
  std::vector<long long> blk_slot;  // input block -> SELL slot (for value updates)
  std::vector<int> sl_off_h;
  int *d_sl_off = nullptr;
  int *d_col = nullptr;
  // the same columns in 16 bits, for matrices of at most 65536 nodes (the
  // step kernel's SpMVs read these when every QP of a launch has them)
  unsigned short *d_col16 = nullptr;
  double *d_val = nullptr;
  double *d_dinv = nullptr;
  int *d_perm = nullptr;
  size_t bytes = 0;
  int version = 0;  // bumped on value updates
  // FP32 copy of the values for the mixed form (made on first use, refreshed on updates)
  float *d_val32 = nullptr;
  int version32 = -1;
  double lambda_min_lower = 0.0;  // > 0: a lower bound on the smallest eigenvalue (error_bound)
};

// Bytes of A one SpMV of the inner solve streams (values and column indices
// of every SELL slot).
inline double matrix_stream_bytes(const mprgp_matrix *A, bool mixed) {
  return (double)A->slots *
         (9.0 * (mixed ? sizeof(float) : sizeof(double)) + (A->d_col16 ? sizeof(unsigned short) : sizeof(int)));
}

struct mprgp_qp {
  mprgp_ctx *ctx = nullptr;
  mprgp_matrix *A = nullptr;
  int n = 0;  // nodes
  double *base = nullptr;
  size_t bytes = 0;
  double *x = nullptr, *x0 = nullptr, *b = nullptr, *g = nullptr, *p = nullptr, *z = nullptr,
         *phi = nullptr, *beta = nullptr, *y = nullptr, *ap = nullptr, *py = nullptr, *ev = nullptr,
         *pv = nullptr, *pw = nullptr, *stage = nullptr, *p2 = nullptr, *beta2 = nullptr, *pu = nullptr,
         *pz = nullptr;
  double *cn = nullptr, *cc = nullptr, *cj = nullptr, *cs = nullptr;
  unsigned char *cf = nullptr, *face = nullptr;
  mprgp::QpState *st = nullptr;
  double *part = nullptr;
  int part_cap = 0;  // tiles
  int m = 0;
  bool has_rhs = false, has_x = false;
  int ev_valid = 0;
  int opcnt[mprgp::OP_NUM] = {};  // device launches per op in the last solve
  // FP32 storage of the mixed form's inner solve (allocated on first use)
  void *base32 = nullptr;
  size_t bytes32 = 0;
  double *fx = nullptr, *fy = nullptr, *fpy = nullptr;  // FP64 iterate/trial points of the inner solve
  float *fb = nullptr, *fg = nullptr, *fphi = nullptr, *fbeta = nullptr;  // FP32 gradient side
  float *fbeta2 = nullptr;
  float *fp = nullptr, *fp2 = nullptr, *fz = nullptr, *fap = nullptr;   // FP32 direction side
  double *fcc = nullptr;  // shifted offsets c - n.x of the correction QP
  std::atomic<int> busy{0};  // submitted to an engine and not yet completed
};

// Rejects calls on a QP an engine is solving.
#define MPRGP_NOT_BUSY(q)                                                        \
  do {                                                                           \
    if ((q)->busy) return ::mprgp::fail(MPRGP_ESTATE, "qp is busy in an engine"); \
  } while (0)
