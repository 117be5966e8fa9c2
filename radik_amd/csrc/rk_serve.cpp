// Serving loop: continuous batching of top-k requests onto the batched
// pointer-form pipeline (the rows entry point).
//
//  * Requests arrive at any time (rk_server_submit, or rk_server_replay, which
//    follows an arrival schedule from the calling thread).  Each request names
//    a device-resident row (values, length, k) and its own output buffers.
//  * One dispatcher thread admits pending requests in arrival order against a
//    device-memory budget: the server's own buffers (per-slot workspace and
//    metadata) plus the input and output bytes of every request in flight.
//  * Admitted requests are coalesced into one batched call on a free slot
//    (its own stream, workspace and metadata).  Up to `streams` batches run at
//    once, so requests that arrive while others are on the GPU start as soon
//    as a slot and budget free up, and their work overlaps on the device.
//  * The batch's metadata is written by the host into mapped pinned memory
//    that the planning kernel reads directly (a host-to-device copy would
//    wait tens of microseconds on the copy engine).  A batch issues the
//    sampled round first, whose last select workgroup copies the workspace
//    header into mapped host memory; only when it reports rows that fell
//    back to the exact passes is the second half of the sequence issued.
//    Completion is detected from a per-slot HIP event polled by the
//    dispatcher.
//  * Latency of a request = host time at which its completion was observed
//    minus its arrival time.
//  * plain = 1 is the baseline the loop is measured against: one request per
//    call through RadiK's exact multi-pass path (RK_EXACT), in arrival order,
//    one stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/radik_amd.h"
#include "rk_pipeline.hpp"

namespace rk {
int topk_rows_part(int dtype, const void* const* row_ptrs, const int64_t* row_len,
                   const int32_t* row_k, int32_t num_rows, int64_t total_len, int32_t max_k,
                   uint32_t flags, void* const* out_val_ptrs, int32_t* const* out_idx_ptrs,
                   void* workspace, size_t ws_bytes, void* stream, int part,
                   uint32_t* publish);
hipError_t publish_header(const void* workspace, uint32_t* dst, hipStream_t stream);
}  // namespace rk

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

enum ReqState : int { RQ_PENDING = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_REJECTED = 3, RQ_FAILED = 4 };

struct Request {
  const void* vals = nullptr;
  int64_t len = 0;
  int32_t k = 0;
  void* out_vals = nullptr;
  int32_t* out_idx = nullptr;
  double t_arrive = 0.0;
  double t_done = 0.0;
  uint64_t bytes = 0;
  int state = RQ_PENDING;
  bool reported = false;  // rk_server_wait / rk_server_replay returned its outcome
};

// Per-row metadata of one batch: the five device arrays the rows entry point
// reads (row pointers, lengths, output pointers, ks), packed in one buffer so
// that one host-to-device copy moves them.
struct Meta {
  const void** ptrs;
  int64_t* lens;
  void** ovals;
  int32_t** oidx;
  int32_t* ks;
};

size_t meta_bytes(int32_t rows) {
  return (size_t)rows * (sizeof(void*) + sizeof(int64_t) + sizeof(void*) + sizeof(int32_t*) +
                         sizeof(int32_t));
}

Meta carve_meta(void* base, int32_t rows) {
  char* p = static_cast<char*>(base);
  Meta m;
  m.ptrs = reinterpret_cast<const void**>(p);
  p += rows * sizeof(void*);
  m.lens = reinterpret_cast<int64_t*>(p);
  p += rows * sizeof(int64_t);
  m.ovals = reinterpret_cast<void**>(p);
  p += rows * sizeof(void*);
  m.oidx = reinterpret_cast<int32_t**>(p);
  p += rows * sizeof(int32_t*);
  m.ks = reinterpret_cast<int32_t*>(p);
  return m;
}

static_assert(offsetof(rk::Header, status) == 0 && offsetof(rk::Header, fallback_rows) == 16,
              "published header words: 0 status, 4 fallback rows");

enum SlotStage : int { SL_IDLE = 0, SL_MAIN = 1, SL_FALLBACK = 2, SL_ALL = 3 };

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* h_meta = nullptr;   // mapped pinned memory: the batch's metadata
  void* d_meta = nullptr;   //   and its device address
  uint32_t* h_hdr = nullptr;  // mapped pinned memory: first 32 bytes of the header
  uint32_t* d_hdr = nullptr;  //   and its device address
  int stage = SL_IDLE;
  // the batch in flight (owned by the dispatcher thread)
  std::vector<uint64_t> tickets;
  int32_t n = 0;
  int64_t keys = 0;
  int32_t max_k = 0;
  uint64_t bytes = 0;
};

}  // namespace

struct rk_server {
  rk_server_config cfg{};
  int device = 0;
  int elem = 4;
  uint32_t flags = 0;
  std::vector<Slot> slots;
  uint64_t server_bytes = 0;

  std::mutex mu;  // guards everything below
  std::condition_variable cv_work, cv_done;
  // Requests still to be reported, ticket t at reqs[t - base]; a request is
  // retired from the front once its outcome has been reported, so memory
  // stays bounded by what is queued, in flight or not yet waited for.
  std::deque<Request> reqs;  // references stay valid while the deque grows
  uint64_t base = 0;
  std::vector<double> latencies_us;  // every completed request, for the statistics
  std::deque<uint64_t> pending;      // tickets not yet admitted, arrival order
  uint64_t inflight = 0;         // request bytes admitted and not completed
  uint64_t peak_inflight = 0;
  int busy_slots = 0;
  bool stop = false;
  int error = RK_OK;
  int64_t batches = 0, batched_rows = 0, fallback_batches = 0, rejected = 0, failed = 0;
  double issue_s = 0.0;  // host time spent issuing batch parts
  int64_t issues = 0;
  int32_t max_rows_seen = 0;

  std::thread dispatcher;

  Request& req(uint64_t ticket) { return reqs[ticket - base]; }
  // Marks a request's outcome as reported and retires reported requests
  // from the front.  Lock held.
  void reported(uint64_t ticket) {
    req(ticket).reported = true;
    while (!reqs.empty() && reqs.front().reported && reqs.front().state >= RQ_DONE) {
      reqs.pop_front();
      ++base;
    }
  }

  void loop();
  bool start_batch(Slot& s);
  int issue(Slot& s, int part);
  void finish(Slot& s, bool ok);
  void set_error(int rc) {
    std::lock_guard<std::mutex> lk(mu);
    if (error == RK_OK) error = rc;
  }
};

// Issues one part of the slot's batch, then the event.  The header reaches
// the slot's mapped host words from the sampled round's last kernel, or from
// a one-wave kernel after the other parts.
int rk_server::issue(Slot& s, int part) {
  const double t0 = now_s();
  Meta d = carve_meta(s.d_meta, cfg.max_batch_rows);
  int rc = rk::topk_rows_part(cfg.dtype, d.ptrs, d.lens, d.ks, s.n, s.keys, std::max(s.max_k, 1),
                              flags, reinterpret_cast<void* const*>(d.ovals),
                              reinterpret_cast<int32_t* const*>(d.oidx), s.ws, s.ws_bytes,
                              s.stream, part, s.d_hdr);
  if (rc != RK_OK) return rc;
  if (part != rk::PART_MAIN && rk::publish_header(s.ws, s.d_hdr, s.stream) != hipSuccess)
    return RK_E_HIP;
  if (hipEventRecord(s.done, s.stream) != hipSuccess) return RK_E_HIP;
  const double dt = now_s() - t0;
  std::lock_guard<std::mutex> lk(mu);
  issue_s += dt;
  ++issues;
  return RK_OK;
}

// Admits requests from the head of the queue into slot `s` and launches them.
// Called by the dispatcher with the lock not held.
bool rk_server::start_batch(Slot& s) {
  std::vector<uint64_t> take;
  int64_t keys = 0;
  uint64_t bytes = 0;
  int32_t max_k = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (pending.empty() || error != RK_OK) return false;
    const uint64_t budget = (uint64_t)cfg.mem_budget - server_bytes;
    while (!pending.empty() && (int32_t)take.size() < cfg.max_batch_rows) {
      Request& r = req(pending.front());
      if (r.bytes > budget) {  // can never be admitted
        r.state = RQ_REJECTED;
        r.t_done = now_s();
        ++rejected;
        pending.pop_front();
        cv_done.notify_all();
        continue;
      }
      if (keys + r.len > cfg.max_batch_keys) break;
      if (inflight + bytes + r.bytes > budget) break;  // wait for memory to free up
      keys += r.len;
      bytes += r.bytes;
      max_k = std::max(max_k, r.k);
      r.state = RQ_RUNNING;
      take.push_back(pending.front());
      pending.pop_front();
    }
    if (take.empty()) return false;
    inflight += bytes;
    peak_inflight = std::max(peak_inflight, inflight);
    ++busy_slots;
    ++batches;
    batched_rows += (int64_t)take.size();
    max_rows_seen = std::max(max_rows_seen, (int32_t)take.size());
    // Copy the admitted requests' fields while the lock is held (the deque
    // may grow concurrently; its elements do not move).
    Meta h = carve_meta(s.h_meta, cfg.max_batch_rows);
    for (size_t i = 0; i < take.size(); ++i) {
      const Request& r = req(take[i]);
      h.ptrs[i] = r.vals;
      h.lens[i] = r.len;
      h.ks[i] = r.k;
      h.ovals[i] = r.out_vals;
      h.oidx[i] = r.out_idx;
    }
  }
  s.n = (int32_t)take.size();
  s.keys = keys;
  s.max_k = max_k;
  s.bytes = bytes;
  s.tickets = std::move(take);
  const int part = cfg.plain ? rk::PART_ALL : rk::PART_MAIN;
  s.stage = cfg.plain ? SL_ALL : SL_MAIN;
  const int rc = issue(s, part);
  if (rc != RK_OK) {
    set_error(rc);
    (void)hipStreamSynchronize(s.stream);
    finish(s, false);
    return false;
  }
  return true;
}

void rk_server::finish(Slot& s, bool ok) {
  const double t = now_s();
  std::lock_guard<std::mutex> lk(mu);
  for (uint64_t tk : s.tickets) {
    Request& r = req(tk);
    r.state = ok ? RQ_DONE : RQ_FAILED;
    r.t_done = t;
    if (ok) latencies_us.push_back((t - r.t_arrive) * 1e6);
  }
  if (!ok) failed += (int64_t)s.tickets.size();
  inflight -= s.bytes;
  --busy_slots;
  s.tickets.clear();
  s.stage = SL_IDLE;
  cv_done.notify_all();
}

void rk_server::loop() {
  (void)hipSetDevice(device);
  double idle_since = now_s();
  for (;;) {
    bool progressed = false;
    int in_flight = 0;
    for (Slot& s : slots) {
      if (s.stage == SL_IDLE) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) {
        ++in_flight;
        continue;
      }
      progressed = true;
      if (q != hipSuccess) {
        set_error(RK_E_HIP);
        finish(s, false);
        continue;
      }
      // words of the published header: 0 status, 4 fallback rows
      const uint32_t status = reinterpret_cast<volatile uint32_t*>(s.h_hdr)[0];
      const uint32_t fallback_rows = reinterpret_cast<volatile uint32_t*>(s.h_hdr)[4];
      if (status != 0) {
        finish(s, false);  // a row failed device-side validation
      } else if (s.stage == SL_MAIN && fallback_rows > 0) {
        // rows that left the sampled path: issue the exact passes
        {
          std::lock_guard<std::mutex> lk(mu);
          ++fallback_batches;
        }
        s.stage = SL_FALLBACK;
        const int rc = issue(s, rk::PART_FALLBACK);
        if (rc != RK_OK) {
          set_error(rc);
          (void)hipStreamSynchronize(s.stream);
          finish(s, false);
        } else {
          ++in_flight;
        }
      } else {
        finish(s, true);
      }
    }
    for (Slot& s : slots) {
      if (s.stage != SL_IDLE) continue;
      if (start_batch(s)) {
        progressed = true;
        ++in_flight;
      }
    }
    bool have_pending;
    {
      std::unique_lock<std::mutex> lk(mu);
      have_pending = !pending.empty();
      if (stop && !have_pending && in_flight == 0) break;
      if (error != RK_OK && in_flight == 0) {
        // after a HIP error: fail what is queued and admit nothing more
        for (uint64_t tk : pending) {
          req(tk).state = RQ_FAILED;
          req(tk).t_done = now_s();
        }
        failed += (int64_t)pending.size();
        pending.clear();
        cv_done.notify_all();
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        continue;
      }
      if (!progressed && in_flight == 0 && !have_pending && now_s() - idle_since > 2e-3) {
        // idle for 2 ms: block until a request arrives
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        idle_since = now_s();
        continue;
      }
    }
    if (progressed || in_flight > 0 || have_pending) idle_since = now_s();
    // While batches are in flight (and shortly after), poll without sleeping:
    // a timed sleep would add tens of microseconds to every completion.
    if (!progressed) std::this_thread::yield();
  }
}

namespace {

void destroy_slots(rk_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.ws) (void)hipFree(sl.ws);
    if (sl.h_meta) (void)hipHostFree(sl.h_meta);
    if (sl.h_hdr) (void)hipHostFree(sl.h_hdr);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  s->slots.clear();
}

}  // namespace

extern "C" {

int rk_server_create(const rk_server_config* cfg, rk_server** out) {
  if (!cfg || !out) return RK_E_ARG;
  *out = nullptr;
  if (cfg->dtype != RK_F32 && cfg->dtype != RK_F16 && cfg->dtype != RK_BF16)
    return RK_E_UNSUPPORTED;
  if (cfg->max_batch_rows <= 0 || cfg->max_batch_keys <= 0 || cfg->streams <= 0 ||
      cfg->mem_budget <= 0 || cfg->max_k <= 0 || (cfg->flags & RK_EXACT))
    return RK_E_ARG;
  if (cfg->max_k > rk_max_k() && (cfg->flags & RK_SORTED)) return RK_E_UNSUPPORTED;
  rk_server* s = new rk_server;
  s->cfg = *cfg;
  if (s->cfg.plain) {  // the baseline: one request per call, one stream
    s->cfg.max_batch_rows = 1;
    s->cfg.streams = 1;
  }
  s->elem = cfg->dtype == RK_F32 ? 4 : 2;
  s->flags = s->cfg.flags | (s->cfg.plain ? (uint32_t)RK_EXACT : 0u);
  if (hipGetDevice(&s->device) != hipSuccess) {
    delete s;
    return RK_E_HIP;
  }
  size_t ws = 0;
  int rc = rk_topk_workspace_size(s->cfg.dtype, s->cfg.max_batch_rows, s->cfg.max_batch_keys,
                                  s->cfg.max_k, s->flags, &ws);
  if (rc != RK_OK) {
    delete s;
    return rc;
  }
  const size_t mb = meta_bytes(s->cfg.max_batch_rows);
  s->slots.resize(s->cfg.streams);
  for (Slot& sl : s->slots) {
    sl.ws_bytes = ws;
    // default-priority, non-blocking streams: batches overlap each other and
    // the caller's work
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&sl.ws, ws) != hipSuccess ||
        hipHostMalloc(&sl.h_meta, mb, hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer(&sl.d_meta, sl.h_meta, 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&sl.h_hdr), 32,
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&sl.d_hdr), sl.h_hdr, 0) !=
            hipSuccess) {
      destroy_slots(s);
      delete s;
      return RK_E_HIP;
    }
    sl.tickets.reserve(s->cfg.max_batch_rows);
    s->server_bytes += ws + mb;
  }
  if (s->server_bytes >= (uint64_t)s->cfg.mem_budget) {
    destroy_slots(s);
    delete s;
    return RK_E_WORKSPACE;
  }
  s->dispatcher = std::thread([s] { s->loop(); });
  *out = s;
  return RK_OK;
}

}  // extern "C"

namespace {

// Queues one request arriving at t_arrive (<= now): the replay passes the
// scheduled arrival, so a late submission does not hide queueing delay.
int submit_at(rk_server* s, const void* vals, int64_t len, int32_t k, void* out_vals,
              int32_t* out_idx, uint64_t* ticket, double t_arrive) {
  if (!s || !ticket || len < 0 || k < 0 || (!vals && len > 0) ||
      (k > 0 && (!out_vals || !out_idx)))
    return RK_E_ARG;
  if (k > s->cfg.max_k || len > s->cfg.max_batch_keys) return RK_E_UNSUPPORTED;
  if ((uintptr_t)vals % s->elem || (uintptr_t)out_vals % s->elem) return RK_E_ALIGN;
  Request r;
  r.vals = vals;
  r.len = len;
  r.k = k;
  r.out_vals = out_vals;
  r.out_idx = out_idx;
  r.bytes = (uint64_t)len * s->elem + (uint64_t)k * (s->elem + sizeof(int32_t));
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error != RK_OK) return s->error;
  r.t_arrive = t_arrive;
  *ticket = s->base + s->reqs.size();
  s->reqs.push_back(r);
  s->pending.push_back(*ticket);
  s->cv_work.notify_one();
  return RK_OK;
}

}  // namespace

extern "C" {

int rk_server_submit(rk_server* s, const void* vals, int64_t len, int32_t k, void* out_vals,
                     int32_t* out_idx, uint64_t* ticket) {
  return submit_at(s, vals, len, k, out_vals, out_idx, ticket, now_s());
}

int rk_server_wait(rk_server* s, uint64_t ticket, double* latency_us) {
  if (!s) return RK_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  // tickets already reported (retired) or never issued
  if (ticket < s->base || ticket >= s->base + s->reqs.size()) return RK_E_ARG;
  s->cv_done.wait(lk, [&] { return s->req(ticket).state >= RQ_DONE; });
  const Request& r = s->req(ticket);
  if (r.reported) return RK_E_ARG;
  if (latency_us) *latency_us = (r.t_done - r.t_arrive) * 1e6;
  const int state = r.state;
  s->reported(ticket);
  switch (state) {
    case RQ_DONE: return RK_OK;
    case RQ_REJECTED: return RK_E_WORKSPACE;
    default: return s->error != RK_OK ? s->error : RK_E_DEVICE;
  }
}

int rk_server_drain(rk_server* s) {
  if (!s) return RK_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->pending.empty() && s->busy_slots == 0; });
  return s->error;
}

int rk_server_replay(rk_server* s, int32_t n, const void* const* vals, const int64_t* lens,
                     const int32_t* ks, void* const* out_vals, int32_t* const* out_idx,
                     const double* arrival_s, double* latency_us, double* makespan_s) {
  if (!s || n < 0 || (n > 0 && (!vals || !lens || !ks || !out_vals || !out_idx || !arrival_s)))
    return RK_E_ARG;
  std::vector<uint64_t> tickets((size_t)n);
  const double t0 = now_s();
  for (int32_t i = 0; i < n; ++i) {
    const double due = t0 + arrival_s[i];
    for (double t = now_s(); t < due; t = now_s()) {
      if (due - t > 3e-4) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    // open-loop load: latency counts from the scheduled arrival
    const int rc = submit_at(s, vals[i], lens[i], ks[i], out_vals[i], out_idx[i], &tickets[i],
                             std::min(now_s(), due));
    if (rc != RK_OK) {
      (void)rk_server_drain(s);
      return rc;
    }
  }
  int rc = rk_server_drain(s);
  std::lock_guard<std::mutex> lk(s->mu);
  double last = t0;
  for (int32_t i = 0; i < n; ++i) {
    const Request& r = s->req(tickets[i]);
    if (latency_us) latency_us[i] = r.state == RQ_DONE ? (r.t_done - r.t_arrive) * 1e6 : -1.0;
    if (r.state == RQ_FAILED && rc == RK_OK) rc = RK_E_DEVICE;
    last = std::max(last, r.t_done);
  }
  for (int32_t i = 0; i < n; ++i) s->reported(tickets[i]);
  if (makespan_s) *makespan_s = last - t0;
  return rc;
}

int rk_server_get_stats(rk_server* s, rk_server_stats* out) {
  if (!s || !out) return RK_E_ARG;
  std::vector<double> lat;
  std::unique_lock<std::mutex> lk(s->mu);
  lat = s->latencies_us;  // a copy under the lock; sorted without it
  std::memset(out, 0, sizeof(*out));
  out->rejected = s->rejected;
  out->failed = s->failed;
  out->batches = s->batches;
  out->fallback_batches = s->fallback_batches;
  out->server_bytes = (int64_t)s->server_bytes;
  out->peak_inflight_bytes = (int64_t)s->peak_inflight;
  out->mean_batch_rows = s->batches ? (double)s->batched_rows / (double)s->batches : 0.0;
  out->max_batch_rows = s->max_rows_seen;
  out->issue_us = s->issues ? s->issue_s / (double)s->issues * 1e6 : 0.0;
  out->retained = (int64_t)s->reqs.size();
  lk.unlock();
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double q) {  // nearest-rank percentile
    if (lat.empty()) return 0.0;
    const size_t rank = (size_t)std::ceil(q * (double)lat.size());
    return lat[std::min(lat.size(), std::max<size_t>(rank, 1)) - 1];
  };
  out->completed = (int64_t)lat.size();
  out->p50_us = pct(0.50);
  out->p99_us = pct(0.99);
  double sum = 0.0;
  for (double v : lat) sum += v;
  out->mean_us = lat.empty() ? 0.0 : sum / (double)lat.size();
  out->max_us = lat.empty() ? 0.0 : lat.back();
  return RK_OK;
}

int rk_server_destroy(rk_server* s) {
  if (!s) return RK_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  destroy_slots(s);
  delete s;
  return RK_OK;
}

}  // extern "C"
