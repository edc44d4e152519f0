// radik_serve.cpp -- serving engine: continuous batching of top-k requests.
//
// submit() queues a request (a rows x cols row-major matrix on the device,
// one task per row, its own k).  A dispatcher thread coalesces pending
// requests, oldest first, into one launch whenever an in-flight slot is
// free and the device budget has room: the budget is one pool shared by all
// slots, charged with each in-flight batch's workspace footprint.  A batch
// takes requests up to the slot share (budget / slots), so a fresh arrival
// waits behind at most one bounded batch per slot; a request larger than
// the share runs alone once the pool has room for it (when the device is
// idle, up to the whole budget).  Each slot owns a stream and a pipeline
// (descriptors, histograms, spill buffers, kept between batches up to the
// share: a slot that ran a larger batch frees them afterwards), so up to
// `slots` batches overlap on the GPU.  One waiter thread per slot blocks on
// the slot's completion event and stamps its requests' latency (host submit
// -> completion) the moment it fires.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/radik.h"
#include "radik_host.h"

using namespace radik;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
  int64_t ticket;
  const void* in;
  int64_t rows, cols, k;
  void* val_out;
  int32_t* idx_out;
  int64_t bytes;  // admission footprint
  Clock::time_point t_submit;
};

struct Result {
  bool done = false;
  double latency_ms = 0;
  Clock::time_point t_submit;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::unique_ptr<Pipeline> pipe;
  std::vector<TaskSpec> specs;
  std::vector<int64_t> tickets;
  int64_t bytes = 0;      // footprint charged to the pool while in flight
  int64_t rows = 0;
  bool busy = false;      // taken by the dispatcher
  bool launched = false;  // its batch and completion event are enqueued
  std::condition_variable cv;  // waiter: batch launched, or stop
  std::thread waiter;
};

}  // namespace

struct radik_server {
  int device = 0;
  int64_t budget = 0;
  int64_t share = 0;   // batch bound: budget / slots
  uint32_t flags = 0;
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_work;   // dispatcher: new request, freed slot or budget
  std::condition_variable cv_done;   // callers of wait/drain: a result completed
  std::deque<Request> pending;
  std::unordered_map<int64_t, Result> results;
  int64_t next_ticket = 1;
  int64_t pool_used = 0;  // footprint of the batches in flight
  bool stop = false;          // no new requests; the dispatcher ends once pending is empty
  bool stop_waiters = false;  // the dispatcher has ended
  int inflight = 0;
  std::thread dispatcher;
  std::atomic<int64_t> launches{0}, rows_done{0}, max_coalesced{0}, max_pool{0};
  int error = 0;
  std::string error_msg;
};

static int64_t footprint(int64_t rows, int64_t cols) {
  // spill buffer + histogram + descriptors per row (upper bound)
  return rows * (cols * 8 + HIST_BINS * 4 + 2 * (int64_t)sizeof(TaskDesc) + (int64_t)sizeof(TaskCtl) + 8);
}

// How many pending requests (footprints bytes[0..n), oldest first) the next
// batch takes: 0 when the oldest does not fit the pool's room (budget -
// pool_used); else it and the following ones while the batch stays within
// the slot share and the room (a first request above the share goes alone).
static int64_t batch_take(const int64_t* bytes, int64_t n, int64_t share, int64_t pool_used, int64_t budget) {
  if (n <= 0 || pool_used + bytes[0] > budget) return 0;
  int64_t used = bytes[0], i = 1;
  while (i < n && used + bytes[i] <= share && pool_used + used + bytes[i] <= budget) used += bytes[i++];
  return i;
}

static void raise_max(std::atomic<int64_t>& a, int64_t v) {
  int64_t prev = a.load();
  while (v > prev && !a.compare_exchange_weak(prev, v)) {
  }
}

static void dispatch_loop(radik_server* s) {
  (void)hipSetDevice(s->device);
  DevFacts* f = nullptr;
  if (dev_facts(&f)) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->error = RADIK_EHIP;
    s->error_msg = "device query failed in dispatcher";
    return;
  }
  const uint32_t tflags = s->flags & 0x0Fu & ~RADIK_WITHIDXIN;
  const int kt = (int)((s->flags & RADIK_KEY_MASK) >> RADIK_KEY_SHIFT);
  const int eb = key_bytes(kt);
  std::vector<int64_t> footprints;
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    // (after stop, requests still pending are dispatched as slots free up)
    s->cv_work.wait(lk, [&] {
      if (s->pending.empty()) return s->stop;
      return s->inflight < (int)s->slots.size() && s->pool_used + s->pending.front().bytes <= s->budget;
    });
    if (s->pending.empty()) return;  // stopping, nothing left
    Slot* slot = nullptr;
    for (auto& sl : s->slots)
      if (!sl->busy) { slot = sl.get(); break; }
    // Coalesce, oldest first, up to the slot share (a larger first request
    // runs alone) and the pool's room.
    footprints.clear();
    for (const Request& r : s->pending) footprints.push_back(r.bytes);
    const int64_t take = batch_take(footprints.data(), (int64_t)footprints.size(), s->share, s->pool_used, s->budget);
    std::vector<Request> batch;
    int64_t used = 0;
    for (int64_t i = 0; i < take; ++i) {
      used += s->pending.front().bytes;
      batch.push_back(s->pending.front());
      s->pending.pop_front();
    }
    slot->busy = true;
    slot->launched = false;
    slot->bytes = used;
    s->pool_used += used;
    s->inflight++;
    raise_max(s->max_pool, s->pool_used);
    lk.unlock();
    slot->specs.clear();
    slot->tickets.clear();
    int64_t nrows = 0;
    for (const Request& r : batch) {
      for (int64_t i = 0; i < r.rows; ++i) {
        TaskSpec sp;
        sp.in = elem_ptr(r.in, i * r.cols, eb);
        sp.idx_in = nullptr;
        sp.val_out = elem_ptr(r.val_out, i * r.k, eb);
        sp.idx_out = r.idx_out + i * r.k;
        sp.len = r.cols;
        sp.k = r.k;
        sp.flags = tflags;
        slot->specs.push_back(sp);
      }
      nrows += r.rows;
      slot->tickets.push_back(r.ticket);
    }
    slot->rows = nrows;
    if (!slot->pipe) slot->pipe.reset(new Pipeline());
    LaunchOpts o;
    o.key_type = kt;
    int rc = stage_tasks(*slot->pipe, slot->specs.data(), (int64_t)slot->specs.size(), *f, o, slot->stream);
    if (!rc) rc = upload(*slot->pipe, slot->stream);
    if (!rc) rc = enqueue(*slot->pipe, *f, o, slot->stream, nullptr);
    if (hipEventRecord(slot->done, slot->stream) != hipSuccess && !rc) rc = RADIK_EHIP;
    s->launches++;
    raise_max(s->max_coalesced, (int64_t)batch.size());
    lk.lock();
    if (rc) {
      s->error = rc;
      s->error_msg = "batch launch failed";
    }
    slot->launched = true;
    slot->cv.notify_all();
  }
}

// One per slot: waits for the slot's batch to complete, stamps its requests
// and returns the slot and its footprint to the pool.
static void waiter_loop(radik_server* s, Slot* slot) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    slot->cv.wait(lk, [&] { return slot->launched || s->stop_waiters; });
    if (!slot->launched) return;  // stopping, no batch left
    lk.unlock();
    const hipError_t e = hipEventSynchronize(slot->done);
    const auto now = Clock::now();
    // A slot that ran a batch above the share frees its buffers, so the
    // slots together never hold more than the budget between batches.
    if (slot->bytes > s->share) slot->pipe.reset();
    lk.lock();
    if (e != hipSuccess && !s->error) {
      s->error = RADIK_EHIP;
      s->error_msg = "a batch failed on the device";
    }
    for (int64_t t : slot->tickets) {
      auto it = s->results.find(t);
      if (it == s->results.end()) continue;
      it->second.done = true;
      it->second.latency_ms = std::chrono::duration<double, std::milli>(now - it->second.t_submit).count();
    }
    s->rows_done += slot->rows;
    s->pool_used -= slot->bytes;
    slot->bytes = 0;
    slot->busy = false;
    slot->launched = false;
    s->inflight--;
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
}

extern "C" {

int radik_server_create(radik_server** out, int64_t budget_bytes, int max_inflight, uint32_t flags) {
  clear_error();
  if (!out) return set_error(RADIK_EINVAL, "srv is NULL");
  *out = nullptr;
  if (budget_bytes <= 0 || max_inflight <= 0 || max_inflight > 16)
    return set_error(RADIK_EINVAL, "budget must be > 0 and max_inflight in 1..16");
  if ((flags & RADIK_WITHSCALE) && (flags & RADIK_KEY_MASK))
    return set_error(RADIK_EINVAL, "WITHSCALE needs Float32 keys");
  if ((int)((flags & RADIK_KEY_MASK) >> RADIK_KEY_SHIFT) >= NUM_KEY_TYPES)
    return set_error(RADIK_EINVAL, "the serving engine takes Float32, Float16, BFloat16 or Int32 keys");
  std::unique_ptr<radik_server> s(new (std::nothrow) radik_server());
  if (!s) return set_error(RADIK_ENOMEM, "out of host memory");
  RADIK_HIP(hipGetDevice(&s->device));
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  s->budget = budget_bytes;
  s->share = budget_bytes / max_inflight;
  s->flags = flags;
  for (int i = 0; i < max_inflight; ++i) {
    std::unique_ptr<Slot> sl(new Slot());
    RADIK_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    RADIK_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
    s->slots.push_back(std::move(sl));
  }
  radik_server* raw = s.get();
  for (auto& sl : raw->slots) sl->waiter = std::thread(waiter_loop, raw, sl.get());
  raw->dispatcher = std::thread(dispatch_loop, raw);
  *out = s.release();
  return 0;
}

int radik_server_submit(radik_server* s, const void* in, int64_t rows, int64_t cols, int64_t k, void* val_out,
                        int32_t* idx_out, int64_t* ticket) {
  clear_error();
  if (!s || !ticket) return set_error(RADIK_EINVAL, "NULL argument");
  if (rows <= 0 || cols <= 0 || k < 0 || k > INT32_MAX || cols > INT32_MAX)
    return set_error(RADIK_EINVAL, "bad request shape rows=%lld cols=%lld k=%lld", (long long)rows,
                     (long long)cols, (long long)k);
  if (!in || (k > 0 && (!val_out || !idx_out))) return set_error(RADIK_EINVAL, "NULL buffer");
  const int64_t bytes = footprint(rows, cols);
  if (bytes > s->budget)
    return set_error(RADIK_EADMIT, "request needs %lld bytes of workspace, the budget is %lld", (long long)bytes,
                     (long long)s->budget);
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return set_error(RADIK_EINVAL, "server is stopping");
  Request r;
  r.ticket = s->next_ticket++;
  r.in = in;
  r.rows = rows;
  r.cols = cols;
  r.k = k;
  r.val_out = val_out;
  r.idx_out = idx_out;
  r.bytes = bytes;
  r.t_submit = Clock::now();
  Result res;
  res.t_submit = r.t_submit;
  s->results[r.ticket] = res;
  s->pending.push_back(r);
  *ticket = r.ticket;
  s->cv_work.notify_all();
  return 0;
}

int radik_server_poll(radik_server* s, int64_t ticket, double* latency_ms) {
  clear_error();
  if (!s) return set_error(RADIK_EINVAL, "srv is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return set_error(RADIK_EINVAL, "unknown ticket %lld", (long long)ticket);
  if (!it->second.done) return RADIK_ENOTREADY;
  if (latency_ms) *latency_ms = it->second.latency_ms;
  s->results.erase(it);
  return 0;
}

int radik_server_wait(radik_server* s, int64_t ticket, double* latency_ms) {
  clear_error();
  if (!s) return set_error(RADIK_EINVAL, "srv is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->results.find(ticket);
  if (it == s->results.end()) return set_error(RADIK_EINVAL, "unknown ticket %lld", (long long)ticket);
  s->cv_done.wait(lk, [&] { return s->error != 0 || s->results[ticket].done; });
  if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
  it = s->results.find(ticket);
  if (latency_ms) *latency_ms = it->second.latency_ms;
  s->results.erase(it);
  return 0;
}

int radik_server_drain(radik_server* s) {
  clear_error();
  if (!s) return set_error(RADIK_EINVAL, "srv is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->error != 0 || (s->pending.empty() && s->inflight == 0); });
  if (s->error) return set_error(s->error, "%s", s->error_msg.c_str());
  return 0;
}

int radik_diag_batch(const int64_t* bytes, int64_t n, int64_t share, int64_t pool_used, int64_t budget,
                     int64_t* take) {
  clear_error();
  if (!take || n < 0 || (n > 0 && !bytes)) return set_error(RADIK_EINVAL, "bad arguments");
  *take = batch_take(bytes, n, share, pool_used, budget);
  return 0;
}

int radik_diag_footprint(int64_t rows, int64_t cols, int64_t* bytes) {
  clear_error();
  if (!bytes || rows < 0 || cols < 0) return set_error(RADIK_EINVAL, "bad arguments");
  *bytes = footprint(rows, cols);
  return 0;
}

int radik_server_stats(radik_server* s, int64_t* launches, int64_t* rows_done, int64_t* max_coalesced) {
  clear_error();
  if (!s) return set_error(RADIK_EINVAL, "srv is NULL");
  if (launches) *launches = s->launches.load();
  if (rows_done) *rows_done = s->rows_done.load();
  if (max_coalesced) *max_coalesced = s->max_coalesced.load();
  return 0;
}

int radik_server_destroy(radik_server* s) {
  clear_error();
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  // the dispatcher launches what is still pending as the waiters free slots
  if (s->dispatcher.joinable()) s->dispatcher.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_waiters = true;
    for (auto& sl : s->slots) sl->cv.notify_all();
  }
  for (auto& sl : s->slots)
    if (sl->waiter.joinable()) sl->waiter.join();
  for (auto& sl : s->slots) {
    if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    if (sl->done) (void)hipEventDestroy(sl->done);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
  }
  delete s;
  return 0;
}

}  // extern "C"
