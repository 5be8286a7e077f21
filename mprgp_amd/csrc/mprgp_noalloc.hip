// mprgp_noalloc.hip -- mprgp_qp_solve_noalloc: the single-QP solve of
// mprgp_qp_solve with no device or host allocation, for callers that cannot
// allocate (tight serving loops, harnesses that forbid hipMalloc).
//
// mprgp_qp_solve keeps its launch tables (QP descriptor, tile table, running
// counter, done flag) in the context's batch tables, which the library
// allocates and grows on demand, and replays captured hipGraphs (whose
// instantiation allocates).  Here those tables live in `scratch`, one
// file-scope device array per GPU of kScratchBytes, and the step kernel is
// launched directly: the same kernel, launch shape and op sequence, so the
// result is bit-identical to mprgp_qp_solve's.  The QP's own vectors, state
// copies and partial sums (made by mprgp_qp_create) are the inputs and
// outputs, and the final state copies come back through the context's pinned
// buffer (made at context creation).  Host code never names `scratch`
// itself: the array's address on each device comes from hipGetSymbolAddress,
// looked up once per device when the first context on it is created.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "mprgp_host.h"

namespace mprgp {

// Workgroup slots the tables hold: QPs of up to 16384 tiles (2^20 nodes at
// one wavefront per workgroup, 2^22 at four).
constexpr int kNoAllocMaxTiles = 16384;

// Scratch layout: control words, the QP descriptor, then the tile table.
constexpr size_t kCtrlOff = 0;          // int[4]: [1] running QPs
constexpr size_t kFlagOff = 64;         // int: the QP's done flag
constexpr size_t kQpDevOff = 256;       // QpDev
constexpr size_t kTileOff = 2048;       // Tile[kNoAllocMaxTiles]
static_assert(sizeof(QpDev) <= kTileOff - kQpDevOff, "QpDev fits its scratch slot");
constexpr size_t kScratchBytes = kTileOff + sizeof(Tile) * (size_t)kNoAllocMaxTiles;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the scratch array's address (set once, when the first context
// on the device is created, with the device current) and a lock that
// serialises the no-alloc solves of this process on that device, which share
// the array.
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices];
static std::mutex g_init_mu;
static std::mutex g_solve_mu[kMaxDevices];

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(MPRGP_ENODEV, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_scratch[device]) return MPRGP_OK;
  void *p = nullptr;
  MPRGP_HIP(hipSetDevice(device));  // the array's copy on this device
  MPRGP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(MPRGP_EHIP, "no address for the no-alloc scratch array on device %d", device);
  g_scratch[device] = p;
  return MPRGP_OK;
}

// Writes the launch tables of one QP into the scratch array and both of the
// QP's state copies: the descriptor, the running count (1), a cleared done
// flag, and the tile table in layout_tiles' XCD order (slot b holds tile
// i = (b/8/c)*8c + (b%8)*c + (b/8)%c, the inverse of its map, or nothing).
// One thread per table slot; thread 0 of block 0 also writes the rest.
struct StateWords {  // a QpState as a kernel argument (plain 8-byte words)
  uint64_t w[64];
};
static_assert(sizeof(QpState) == sizeof(StateWords), "a state copy is 64 words");

__global__ void __launch_bounds__(256) noalloc_tables(unsigned char *base, QpDev d, StateWords st, QpState *qst,
                                                      double *part, int n_slices, int span, int n_tiles,
                                                      int grid, int run) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b < grid) {
    Tile t{nullptr, nullptr, -1, 0, 0, 0, 0, 0};
    const int xcd = b % 8, pos = b / 8;
    const int i = (pos / run) * (8 * run) + xcd * run + pos % run;
    if (i < n_tiles) {
      const int s0 = i * span;
      t = Tile{qst, part, 0, s0, min(span, n_slices - s0), i, n_tiles, 0};
    }
    reinterpret_cast<Tile *>(base + kTileOff)[b] = t;
  }
  if (b == 0) {
    *reinterpret_cast<QpDev *>(base + kQpDevOff) = d;
    int *ctrl = reinterpret_cast<int *>(base + kCtrlOff);
    ctrl[0] = 0;
    ctrl[1] = 1;
    *reinterpret_cast<int *>(base + kFlagOff) = 0;
  }
  if (b < 2 * 64) reinterpret_cast<uint64_t *>(qst)[b] = st.w[b % 64];  // the two state copies
}

}  // namespace mprgp

using namespace mprgp;

extern "C" int mprgp_qp_solve_noalloc(mprgp_qp *q, const mprgp_params *params, mprgp_stats *stats) {
  MPRGP_TRY
  if (!q) return fail(MPRGP_EBADINPUT, "NULL qp");
  mprgp_params def;
  if (!params) {
    mprgp_params_default(&def);
    params = &def;
  }
  const mprgp_params &p = *params;
  mprgp_ctx *ctx = q->ctx;
  if (int rc = check_batch_qp(ctx, q, p, 0)) return rc;
  if (p.precision == MPRGP_MIXED && (!q->base32 || !q->A->d_val32))
    return fail(MPRGP_ESTATE, "the mixed form's FP32 storage is made by a first mprgp_qp_solve of this QP; "
                              "the no-alloc solve does not allocate it");
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch[dev])
    return fail(MPRGP_ESTATE, "no scratch array address for device %d", dev);
  const int ts = pick_ts(p, q->A->n_slices);
  const int spw = slices_per_wave_of(p);
  const int span = ts * spw;  // as mprgp_qp_solve
  const bool c16 = use_col16(q->A->d_col16 != nullptr, spw);
  const int n_tiles = tiles_for(q->A->n_slices, span);
  const int grid = round_grid(n_tiles);
  if (grid > kNoAllocMaxTiles)
    return fail(MPRGP_EBADINPUT, "the QP needs %d workgroup slots of %d wavefronts, the no-alloc tables hold %d",
                grid, ts, kNoAllocMaxTiles);
  MPRGP_HIP(hipSetDevice(dev));
  std::lock_guard<std::recursive_mutex> lk_(ctx->mu);
  std::lock_guard<std::mutex> lk_dev(g_solve_mu[dev]);
  unsigned char *base = static_cast<unsigned char *>(g_scratch[dev]);
  hipStream_t s = ctx->stream;
  const bool mixed = p.precision == MPRGP_MIXED;
  QpDev d;
  QpState st;
  describe_qp(q, p, n_tiles, reinterpret_cast<int *>(base + kFlagOff), 1, d, st);
  if (p.from_x0) MPRGP_HIP(hipMemcpyAsync(q->x, q->x0, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, s));
  // the QP's set-up, as prepare_qp: FP32 values if the matrix changed, the
  // power start vector, the rows' tangent-Jacobi factors
  if (mixed && q->A->version32 != q->A->version) {
    MPRGP_HIP(launch_to_f32(s, q->A->d_val, q->A->d_val32, (size_t)q->A->slots * 9));
    q->A->version32 = q->A->version;
  }
  if (!(p.alpha_bar > 0.0)) {
    if (p.use_ev && q->ev_valid) {
      MPRGP_HIP(hipMemcpyAsync(q->pv, q->ev, sizeof(double) * 3 * q->n, hipMemcpyDeviceToDevice, s));
    } else {
      const size_t n3 = 3 * (size_t)q->n;
      MPRGP_HIP(launch_power_start(s, q->stage, n3, power_start_norm(n3)));
      MPRGP_HIP(launch_gather_nodes(s, q->stage, q->pv, q->A->d_perm, q->n));
    }
  }
  MPRGP_HIP(launch_setup_rows(s, q->A->d_dinv, q->cn, q->cf, q->cs, q->n));
  const int run = std::min(4, grid / 8);  // layout_tiles' runs of tiles per XCD
  StateWords sw;
  std::memcpy(&sw, &st, sizeof(sw));
  noalloc_tables<<<(grid + 255) / 256, 256, 0, s>>>(base, d, sw, q->st, q->part, q->A->n_slices, span, n_tiles,
                                                    grid, run);
  MPRGP_HIP(hipGetLastError());
  int steps = p.graph_steps > 0 ? p.graph_steps : 16;
  if (p.graph_steps < 0) steps = -p.graph_steps;
  steps += steps & 1;
  const long max_steps = solve_step_bound(p);
  const QpDev *qps = reinterpret_cast<const QpDev *>(base + kQpDevOff);
  const Tile *tiles = reinterpret_cast<const Tile *>(base + kTileOff);
  int *running = reinterpret_cast<int *>(base + kCtrlOff) + 1;
  const int variant = step_variant(n_tiles, matrix_stream_bytes(q->A, mixed), tail_tiles_of(p));
  MPRGP_HIP(hipEventRecord(ctx->t0, s));
  long launched = 0;
  for (;;) {
    for (int i = 0; i < steps; ++i)
      MPRGP_HIP(launch_step(ts, grid, s, qps, tiles, running, i & 1, mixed, variant, spw, c16));
    launched += steps;
    MPRGP_HIP(hipMemcpyAsync(ctx->h_running, running, sizeof(int), hipMemcpyDeviceToHost, s));
    MPRGP_HIP(hipStreamSynchronize(s));
    if (ctx->h_running[0] == 0) break;
    if (launched > max_steps + 2L * steps)
      return fail(MPRGP_EHIP, "solve did not finish within %ld device steps", launched);
  }
  MPRGP_HIP(hipEventRecord(ctx->t1, s));
  QpState *both = static_cast<QpState *>(ctx->h_final);
  MPRGP_HIP(hipMemcpyAsync(both, q->st, 2 * sizeof(QpState), hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  MPRGP_HIP(hipEventElapsedTime(&ms, ctx->t0, ctx->t1));
  return fill_stats(q, both, ms, launched, stats);
  MPRGP_CATCH
}
