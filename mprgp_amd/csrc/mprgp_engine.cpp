// mprgp_engine.cpp -- continuous-batching serving engine.
//
// A worker thread owns one HIP stream and one set of batch tables and runs
// "rounds".  Each round:
//   1. applies table changes: QPs found finished in an earlier poll leave the
//      tile table (their two state copies are copied out), queued solves are
//      admitted into free slots while the slot count and the HBM budget allow;
//   2. replays a captured graph of `steps` mprgp_step launches over the table;
//   3. copies the per-slot done tags to pinned memory and records an event;
//   4. waits for the previous round's event and reads its tags and states.
// Every upload goes through pinned staging owned by the round parity, so the
// host never drains the stream: table changes are ordered between replays by
// the stream itself.  Finished QPs are detected by the tag the kernel writes
// into their slot's flag, so a reused slot never needs clearing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "mprgp_host.h"

using namespace mprgp;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
  long long ticket = 0;
  mprgp_qp *q = nullptr;
  mprgp_params p{};
  size_t bytes = 0;
  Clock::time_point t_submit, t_admit;
  double latency_ms = 0.0, queue_ms = 0.0;
  long launches_at_admit = 0;
  mprgp_stats stats{};
  int rc = MPRGP_OK;
  bool done = false;
};

struct Slot {
  Request *req = nullptr;
  int tag = 0;
  long admit_round = 0;
  int n_tiles = 0;
  int retiring = 0;         // 1: leaves the table next round; 2: state copy in flight
  long retire_round = -1;   // round whose event covers the state copy
  QpState *h_state = nullptr;  // pinned destination of the state copy
};

// Pinned host buffer, grown on demand (after the stream drained).
struct Pinned {
  void *p = nullptr;
  size_t cap = 0;
  int reserve(size_t bytes, hipStream_t s) {
    if (bytes <= cap) return MPRGP_OK;
    MPRGP_HIP(hipStreamSynchronize(s));
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t c = std::max(bytes, (size_t)4096);
    MPRGP_HIP(hipHostMalloc(&p, c, hipHostMallocDefault));
    cap = c;
    return MPRGP_OK;
  }
  void free() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace

struct mprgp_engine {
  mprgp_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  BatchTables tab;
  int max_active = 256;
  int ts = 4;
  int steps = 8;
  size_t budget = 0;

  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<long long, Request *> reqs;
  std::deque<long long> done_q;  // completed, not yet collected, in completion order
  long long next_ticket = 1;
  bool stop = false;
  int fatal = MPRGP_OK;  // worker error: every pending ticket fails with it
  std::string fatal_msg;

  // worker-owned
  std::vector<Slot> slots;
  int next_tag = 1;
  size_t active_bytes = 0, peak_bytes = 0;
  int active = 0, peak_active = 0;
  long long submitted = 0, completed = 0, rounds = 0, launches = 0;
  Pinned h_qps[2], h_tiles[2], h_states_in[2], h_flags[2], h_states_out[2];
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<QpDev> devs;  // host mirror of the descriptor table

  void run();
  int round_once(long r, bool &idle);
  void finish(Request *rq, int rc);
};

void mprgp_engine::finish(Request *rq, int rc) {
  rq->rc = rc;
  rq->latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - rq->t_submit).count();
  rq->stats.elapsed_ms = (float)rq->latency_ms;
  rq->stats.queue_ms = rq->queue_ms;
  rq->done = true;
  done_q.push_back(rq->ticket);
  completed++;
}

// One round (see the file comment).  Sets `idle` when nothing is admitted,
// queued or in flight.
int mprgp_engine::round_once(long r, bool &idle) {
  const int par = (int)(r & 1);
  hipStream_t s = stream;
  idle = false;
  // 1a. QPs found finished last round leave the table; their states are copied out
  bool changed = false;
  QpState *st_out = (QpState *)h_states_out[par].p;
  int n_out = 0;
  for (Slot &sl : slots) {
    if (sl.req && sl.retiring == 1) {
      sl.h_state = st_out + 2 * n_out++;
      MPRGP_HIP(hipMemcpyAsync(sl.h_state, sl.req->q->st, 2 * sizeof(QpState), hipMemcpyDeviceToHost, s));
      sl.retiring = 2;
      sl.retire_round = r;
      changed = true;
    }
  }
  // 1b. admissions
  std::vector<Request *> admit;
  {
    std::lock_guard<std::mutex> lk(mu);
    int free_slots = 0;
    for (Slot &sl : slots) free_slots += sl.req == nullptr;
    while (!queue.empty() && free_slots > 0) {
      Request *rq = queue.front();
      const bool fits = budget == 0 || active == 0 || active_bytes + rq->bytes <= budget;
      if (!fits) break;  // FIFO: later requests wait behind the head
      queue.pop_front();
      admit.push_back(rq);
      active_bytes += rq->bytes;
      active++;
      free_slots--;
    }
    peak_active = std::max(peak_active, active);
    peak_bytes = std::max(peak_bytes, active_bytes);
  }
  QpState *st_in = (QpState *)h_states_in[par].p;
  int n_in = 0;
  for (Request *rq : admit) {
    int si = 0;
    while (slots[si].req) ++si;
    Slot &sl = slots[si];
    sl.req = rq;
    sl.tag = next_tag++;
    if (next_tag > 0x3fffffff) next_tag = 1;
    sl.admit_round = r;
    sl.retiring = 0;
    sl.retire_round = -1;
    sl.n_tiles = tiles_for(rq->q->A->n_slices, ts);
    rq->t_admit = Clock::now();
    rq->queue_ms = std::chrono::duration<double, std::milli>(rq->t_admit - rq->t_submit).count();
    rq->launches_at_admit = launches;
    QpState *pst = st_in + 2 * n_in++;
    int rc = prepare_qp(rq->q, rq->p, s, sl.n_tiles, tab.d_flags + si, sl.tag, devs[si], pst[0]);
    if (rc) return rc;
    pst[1] = pst[0];
    MPRGP_HIP(hipMemcpyAsync(rq->q->st, pst, 2 * sizeof(QpState), hipMemcpyHostToDevice, s));
    changed = true;
  }
  // 1c. table upload
  int n_tiles = 0;
  bool any_mixed = false, all_c16 = true;
  double matrix_bytes = 0.0;
  for (const Slot &sl : slots)
    if (sl.req && sl.retiring == 0) {
      n_tiles += sl.n_tiles;
      any_mixed |= sl.req->p.precision == MPRGP_MIXED;
      all_c16 &= sl.req->q->A->d_col16 != nullptr;
      matrix_bytes += matrix_stream_bytes(sl.req->q->A, sl.req->p.precision == MPRGP_MIXED);
    }
  const int grid = round_grid(n_tiles);
  if (changed) {
    if (int rc = h_tiles[par].reserve(sizeof(Tile) * (size_t)grid, s)) return rc;
    // a larger tile table frees the old one and the graphs that captured it:
    // let the replays in flight finish first
    if (grid > tab.tiles_cap) MPRGP_HIP(hipStreamSynchronize(s));
    if (int rc = tab.ensure(max_active, grid)) return rc;
    Tile *tl = (Tile *)h_tiles[par].p;
    std::vector<Tile> tv;
    tv.reserve(n_tiles);
    for (int i = 0; i < (int)slots.size(); ++i)
      if (slots[i].req && slots[i].retiring == 0) append_tiles(i, slots[i].req->q, ts, tv);
    layout_tiles(tv, grid, tl);
    std::memcpy(h_qps[par].p, devs.data(), sizeof(QpDev) * devs.size());
    MPRGP_HIP(hipMemcpyAsync(tab.d_qps, h_qps[par].p, sizeof(QpDev) * devs.size(), hipMemcpyHostToDevice, s));
    MPRGP_HIP(hipMemcpyAsync(tab.d_tiles, tl, sizeof(Tile) * grid, hipMemcpyHostToDevice, s));
  }
  // 2. replay
  if (n_tiles > 0) {
    hipGraphExec_t ex = nullptr;
    if (int rc = tab.graph(s, ts, steps, grid, any_mixed, step_variant(n_tiles, matrix_bytes), &ex, 1,
                           use_col16(all_c16, 1)))
      return rc;
    MPRGP_HIP(hipGraphLaunch(ex, s));
    launches += steps;
  }
  // 3. poll copy
  MPRGP_HIP(hipMemcpyAsync(h_flags[par].p, tab.d_flags, sizeof(int) * (size_t)max_active, hipMemcpyDeviceToHost, s));
  MPRGP_HIP(hipEventRecord(ev[par], s));
  rounds++;
  // 4. previous round's results
  if (r == 0) return MPRGP_OK;
  const int pp = (int)((r - 1) & 1);
  MPRGP_HIP(hipEventSynchronize(ev[pp]));
  const int *flags = (const int *)h_flags[pp].p;
  std::vector<Request *> finished;
  for (int i = 0; i < (int)slots.size(); ++i) {
    Slot &sl = slots[i];
    if (!sl.req) continue;
    if (sl.retiring == 2 && sl.retire_round == r - 1) {
      Request *rq = sl.req;
      int rc = fill_stats(rq->q, sl.h_state, 0.f, launches - rq->launches_at_admit, &rq->stats);
      if (rc == MPRGP_EHIP) set_error("engine: QP left the batch before reaching DONE");
      finished.push_back(rq);
      rq->rc = rc;
      sl.req = nullptr;
      sl.retiring = 0;
      std::memset(&devs[i], 0, sizeof(QpDev));
    } else if (sl.retiring == 0 && sl.admit_round <= r - 1 && flags[i] == sl.tag) {
      sl.retiring = 1;
    }
  }
  if (!finished.empty()) {
    std::lock_guard<std::mutex> lk(mu);
    for (Request *rq : finished) {
      active_bytes -= rq->bytes;
      active--;
      finish(rq, rq->rc);
    }
    cv_done.notify_all();
  }
  bool any = false;
  for (const Slot &sl : slots) any |= sl.req != nullptr;
  std::lock_guard<std::mutex> lk(mu);
  idle = !any && queue.empty();
  return MPRGP_OK;
}

void mprgp_engine::run() {
  (void)hipSetDevice(ctx->device);
  long r = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      bool any = false;
      for (const Slot &sl : slots) any |= sl.req != nullptr;
      if (!any && queue.empty()) {
        if (stop) break;
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        if (stop && queue.empty()) break;
        r = 0;  // the stream is idle: restart the two-deep pipeline
      }
    }
    bool idle = false;
    int rc;
    try {
      rc = round_once(r, idle);
    } catch (...) {  // e.g. std::bad_alloc: the worker fails its tickets instead of terminating
      rc = on_exception();
    }
    if (rc) {
      // fail every pending ticket with the error; the engine stops accepting work
      std::lock_guard<std::mutex> lk(mu);
      fatal = rc;
      fatal_msg = mprgp_last_error();
      (void)hipStreamSynchronize(stream);
      for (Slot &sl : slots)
        if (sl.req) {
          finish(sl.req, rc);
          sl.req = nullptr;
        }
      for (Request *rq : queue) finish(rq, rc);
      queue.clear();
      active = 0;
      active_bytes = 0;
      cv_done.notify_all();
      break;
    }
    ++r;
    if (idle) {
      // drain the last round's event so the next burst starts a fresh pipeline
      (void)hipEventSynchronize(ev[(r - 1) & 1]);
      r = 0;
    }
  }
}

extern "C" void mprgp_engine_config_default(mprgp_engine_config *cfg) {
  MPRGP_TRY
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->max_active = 256;
  cfg->tiles = 4;
  cfg->graph_steps = 8;
  MPRGP_CATCH_VOID
}

extern "C" int mprgp_engine_create(mprgp_ctx *ctx, const mprgp_engine_config *cfg, mprgp_engine **out) {
  MPRGP_TRY
  if (!ctx || !out) return fail(MPRGP_EBADINPUT, "NULL argument");
  *out = nullptr;
  mprgp_engine_config c;
  mprgp_engine_config_default(&c);
  if (cfg) {
    if (cfg->max_active > 0) c.max_active = cfg->max_active;
    if (cfg->tiles) c.tiles = cfg->tiles;
    if (cfg->graph_steps > 0) c.graph_steps = cfg->graph_steps;
    c.hbm_budget = cfg->hbm_budget;
  }
  if (c.tiles != 1 && c.tiles != 2 && c.tiles != 4 && c.tiles != 8)
    return fail(MPRGP_EBADINPUT, "tiles must be 1, 2, 4 or 8");
  if (c.max_active > 65536) return fail(MPRGP_EBADINPUT, "max_active too large");
  MPRGP_HIP(hipSetDevice(ctx->device));
  mprgp_engine *e = new mprgp_engine();
  e->ctx = ctx;
  e->max_active = c.max_active;
  e->ts = c.tiles;
  e->steps = c.graph_steps + (c.graph_steps & 1);
  e->budget = c.hbm_budget;
  e->slots.resize(e->max_active);
  e->devs.assign(e->max_active, QpDev{});
  int rc = e->tab.init(ctx);
  hipError_t he = hipSuccess;
  if (!rc) he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  for (int i = 0; i < 2 && !rc && he == hipSuccess; ++i) he = hipEventCreateWithFlags(&e->ev[i], hipEventDisableTiming);
  if (!rc && he != hipSuccess) rc = fail(MPRGP_EHIP, "engine setup: %s", hipGetErrorString(he));
  for (int i = 0; i < 2 && !rc; ++i) {
    if (!rc) rc = e->h_qps[i].reserve(sizeof(QpDev) * (size_t)e->max_active, e->stream);
    if (!rc) rc = e->h_states_in[i].reserve(2 * sizeof(QpState) * (size_t)e->max_active, e->stream);
    if (!rc) rc = e->h_states_out[i].reserve(2 * sizeof(QpState) * (size_t)e->max_active, e->stream);
    if (!rc) rc = e->h_flags[i].reserve(sizeof(int) * (size_t)e->max_active, e->stream);
    if (!rc) rc = e->h_tiles[i].reserve(sizeof(Tile) * 1024, e->stream);
  }
  if (!rc) rc = e->tab.ensure(e->max_active, 1024);
  if (!rc) {
    hipError_t m = hipMemsetAsync(e->tab.d_flags, 0, sizeof(int) * (size_t)e->max_active, e->stream);
    if (m == hipSuccess) m = hipStreamSynchronize(e->stream);
    if (m != hipSuccess) rc = fail(MPRGP_EHIP, "engine setup: %s", hipGetErrorString(m));
  }
  if (rc) {
    e->stop = true;
    mprgp_engine_destroy(e);
    return rc;
  }
  e->worker = std::thread([e] { e->run(); });
  *out = e;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" int mprgp_engine_submit(mprgp_engine *e, mprgp_qp *q, const mprgp_params *params, long long *ticket) {
  MPRGP_TRY
  if (!e || !q || !ticket) return fail(MPRGP_EBADINPUT, "NULL argument");
  *ticket = 0;
  mprgp_params def;
  if (!params) {
    mprgp_params_default(&def);
    params = &def;
  }
  // QPs of any context on the engine's device (each session may keep its own
  // context, so its uploads do not serialise with other sessions')
  if (q->ctx->device != e->ctx->device) return fail(MPRGP_EBADINPUT, "qp lives on another device");
  if (int rc = check_batch_qp(q->ctx, q, *params, 0)) return rc;
  const size_t bytes = solve_bytes(q, *params);
  if (e->budget && bytes > e->budget)
    return fail(MPRGP_EBUDGET, "QP needs %zu bytes, engine budget is %zu", bytes, e->budget);
  int expected = 0;
  if (!q->busy.compare_exchange_strong(expected, 1)) return fail(MPRGP_ESTATE, "qp is already submitted");
  Request *rq = new Request();
  rq->q = q;
  rq->p = *params;
  rq->bytes = bytes;
  rq->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->fatal || e->stop) {
      q->busy = 0;
      delete rq;
      return e->fatal ? fail(e->fatal, "engine stopped: %s", e->fatal_msg.c_str())
                      : fail(MPRGP_ESTATE, "engine is stopping");
    }
    rq->ticket = e->next_ticket++;
    e->reqs[rq->ticket] = rq;
    e->queue.push_back(rq);
    e->submitted++;
    *ticket = rq->ticket;
  }
  e->cv_work.notify_one();
  return MPRGP_OK;
  MPRGP_CATCH
}

// Hands a completed request's result to the caller and forgets it (engine
// mutex held).
static int collect(mprgp_engine *e, Request *rq, mprgp_stats *stats) {
  if (stats) *stats = rq->stats;
  const int rc = rq->rc;
  rq->q->busy = 0;
  auto d = std::find(e->done_q.begin(), e->done_q.end(), rq->ticket);
  if (d != e->done_q.end()) e->done_q.erase(d);
  e->reqs.erase(rq->ticket);
  delete rq;
  if (rc == MPRGP_EHIP && e->fatal) return fail(rc, "engine failed: %s", e->fatal_msg.c_str());
  return rc;
}

extern "C" int mprgp_engine_wait(mprgp_engine *e, long long ticket, int timeout_ms, mprgp_stats *stats) {
  MPRGP_TRY
  if (!e) return fail(MPRGP_EBADINPUT, "NULL engine");
  std::unique_lock<std::mutex> lk(e->mu);
  auto it = e->reqs.find(ticket);
  if (it == e->reqs.end()) return fail(MPRGP_EBADINPUT, "unknown or already collected ticket %lld", ticket);
  Request *rq = it->second;
  if (timeout_ms < 0) {
    e->cv_done.wait(lk, [&] { return rq->done; });
  } else if (!e->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return rq->done; })) {
    return MPRGP_EPENDING;
  }
  return collect(e, rq, stats);
  MPRGP_CATCH
}

extern "C" int mprgp_engine_wait_any(mprgp_engine *e, int timeout_ms, long long *ticket, mprgp_stats *stats) {
  MPRGP_TRY
  if (!e || !ticket) return fail(MPRGP_EBADINPUT, "NULL argument");
  *ticket = 0;
  std::unique_lock<std::mutex> lk(e->mu);
  auto ready = [&] { return !e->done_q.empty(); };
  if (timeout_ms < 0) {
    if (e->reqs.empty()) return fail(MPRGP_ESTATE, "no ticket outstanding");
    e->cv_done.wait(lk, ready);
  } else if (!e->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return MPRGP_EPENDING;
  }
  auto it = e->reqs.find(e->done_q.front());
  *ticket = it->first;
  return collect(e, it->second, stats);
  MPRGP_CATCH
}

extern "C" int mprgp_engine_get_info(mprgp_engine *e, mprgp_engine_info *info) {
  MPRGP_TRY
  if (!e || !info) return fail(MPRGP_EBADINPUT, "NULL argument");
  std::lock_guard<std::mutex> lk(e->mu);
  std::memset(info, 0, sizeof(*info));
  info->submitted = e->submitted;
  info->completed = e->completed;
  info->rounds = e->rounds;
  info->launches = e->launches;
  info->active = e->active;
  info->queued = (int)e->queue.size();
  info->peak_active = e->peak_active;
  info->active_bytes = e->active_bytes;
  info->peak_bytes = e->peak_bytes;
  return MPRGP_OK;
  MPRGP_CATCH
}

extern "C" void mprgp_engine_destroy(mprgp_engine *e) {
  MPRGP_TRY
  if (!e) return;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->stop = true;
  }
  e->cv_work.notify_all();
  if (e->worker.joinable()) e->worker.join();
  (void)hipSetDevice(e->ctx->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (auto &kv : e->reqs) {
    if (kv.second->q) kv.second->q->busy = 0;
    delete kv.second;
  }
  e->tab.destroy();
  for (int i = 0; i < 2; ++i) {
    e->h_qps[i].free();
    e->h_tiles[i].free();
    e->h_states_in[i].free();
    e->h_states_out[i].free();
    e->h_flags[i].free();
    if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  MPRGP_CATCH_VOID
}
