// radik_api.cpp -- the extern "C" surface of libradik_hip.so: workspace,
// drop-in top-k, batched plans (hipGraph replay), diagnostics and the
// device-memory helpers the ctypes harness uses.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/radik.h"
#include "../../include/radik_devmem.h"
#include "radik_host.h"

namespace radik {
const char* last_error();
int diag_scan(const uint32_t* in, uint32_t* ow, uint32_t* ob, int64_t n, hipStream_t s);
int diag_pass1(Pipeline& p, const DevFacts& f, uint32_t* hist_out, uint32_t* sel_out, hipStream_t s);
int diag_copy(void* dst, const void* src, int64_t bytes, hipStream_t s);
int sample_topp(const void* vals, const void* idx, int idx_bytes, const int64_t* out_off, const int32_t* m,
                int64_t num_tasks, int key_type, float temperature, float top_p, const float* u, int64_t* tokens,
                int32_t* ranks, hipStream_t s);
int64_t diag_plan_chunks(int64_t total, int64_t blocks, int64_t spread, int guided, int64_t gran, int64_t min_chunk,
                         int64_t* out, int64_t cap);
void diag_route(const int64_t* lens, const int64_t* ks, int64_t n, uint32_t small_cap, int predict, int medium,
                int32_t* cls, int64_t* groups);
int noalloc_run(float* val_out, void* idx_out, int idx_bytes, const float* val_in, const void* idx_in,
                const int64_t* task_lens, int64_t num_tasks, int64_t k, uint32_t tflags, hipStream_t s);
}  // namespace radik

using namespace radik;

struct radik_ws {
  int device = 0;
  int64_t n = 0;
  int64_t num_tasks = 0;
  int counter_bytes = 4;
  Pipeline pipe;
  std::vector<TaskSpec> specs;
  hipStream_t last_stream = nullptr;  // the stream of the last call (its work may still reference the buffers)
  bool used = false;
  F64Pipe* f64 = nullptr;  // buffers of Float64 calls (made on the first one)
  ~radik_ws() {
    if (f64) f64_destroy(f64);
  }
};

struct radik_plan {
  int device = 0;
  std::vector<int64_t> lens;
  std::vector<int64_t> ks;
  std::vector<int64_t> in_off, out_off;
  uint32_t flags = 0;
  int idx_bytes = 4;
  Pipeline pipe;
  LaunchOpts opts;
  std::vector<TaskSpec> specs;
  // pointers the staged descriptors were built for
  const void* last_in = nullptr;
  const void* last_idx_in = nullptr;
  void* last_val_out = nullptr;
  void* last_idx_out = nullptr;
  bool staged = false;
  // graph
  bool use_graph = false;  // eager launches (measured faster than graph replay on MI355X)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  // timing
  PhaseTimer timer;
  double ms[NUM_PHASES] = {};
  int64_t timed = 0;
  // sampling (radik_plan_sample): per-task output offsets and result counts
  int64_t* d_soff = nullptr;
  int32_t* d_sm = nullptr;
  ~radik_plan() {
    if (d_soff) (void)hipFree(d_soff);
    if (d_sm) (void)hipFree(d_sm);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : timer.ev)
      if (e) (void)hipEventDestroy(e);
  }
};

// One cached workspace per device for radik_topk (the allocating drop-in).
constexpr int kMaxDevices = 64;
struct TopkCache {
  std::mutex mu;
  radik_ws* ws = nullptr;
};
static TopkCache g_topk_cache[kMaxDevices];

static const uint32_t kFlagMask =
    RADIK_LARGEST | RADIK_ASCEND | RADIK_WITHSCALE | RADIK_WITHIDXIN | RADIK_WITHPACKING | RADIK_KEY_MASK;

static inline int key_type_of(uint32_t flags) { return (int)((flags & RADIK_KEY_MASK) >> RADIK_KEY_SHIFT); }

// WITHSCALE subtracts in Float32 (utils.jl:62-76); other key types reject it.
static int check_key_flags(uint32_t flags) {
  const int kt = key_type_of(flags);
  if (kt >= NUM_KEY_TYPES && kt != KEY_F64) return set_error(RADIK_EINVAL, "unknown key type %d", kt);
  if ((flags & RADIK_WITHSCALE) && kt != KEY_F32) return set_error(RADIK_EINVAL, "WITHSCALE needs Float32 keys");
  return 0;
}

static LaunchOpts opts_for(uint32_t flags, int vload_bytes) {
  LaunchOpts o;
  o.key_type = key_type_of(flags);
  if (!(flags & RADIK_WITHPACKING)) o.unr = 1;
  else if (vload_bytes >= 64) o.unr = 4;
  else if (vload_bytes >= 32) o.unr = 2;
  else o.unr = 1;
  return o;
}

static int check_lens(const int64_t* lens, int64_t T, int idx_bytes, int64_t* total, int64_t* maxlen) {
  int64_t tot = 0, mx = 0;
  for (int64_t t = 0; t < T; ++t) {
    const int64_t l = lens[t];
    if (l < 0) return set_error(RADIK_EINVAL, "task_lens[%lld] = %lld is negative", (long long)t, (long long)l);
    const int64_t lim = (idx_bytes == 4) ? (int64_t)INT32_MAX : (int64_t)UINT32_MAX;
    if (l > lim)
      return set_error(RADIK_EINVAL, "task_lens[%lld] = %lld exceeds %lld elements per task",
                       (long long)t, (long long)l, (long long)lim);
    tot += l;
    if (l > mx) mx = l;
  }
  *total = tot;
  *maxlen = mx;
  return 0;
}

extern "C" {

const char* radik_version(void) { return "radik-mi355x 0.1.0 (gfx950)"; }

const char* radik_last_error(void) { return last_error(); }

// ---------------------------------------------------------------------------
// Workspace and drop-in
// ---------------------------------------------------------------------------
int radik_ws_create(radik_ws** ws, int64_t n, int64_t num_tasks, int counter_bytes, void* stream) {
  clear_error();
  if (!ws) return set_error(RADIK_EINVAL, "ws is NULL");
  *ws = nullptr;
  if (n < 0 || num_tasks < 0) return set_error(RADIK_EINVAL, "n and num_tasks must be >= 0");
  if (counter_bytes != 4 && counter_bytes != 8)
    return set_error(RADIK_EINVAL, "counter_bytes must be 4 (Int32) or 8 (Int64), got %d", counter_bytes);
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  std::unique_ptr<radik_ws> w(new (std::nothrow) radik_ws());
  if (!w) return set_error(RADIK_ENOMEM, "out of host memory");
  RADIK_HIP(hipGetDevice(&w->device));
  w->n = n;
  w->num_tasks = num_tasks;
  w->counter_bytes = counter_bytes;
  // Pre-size descriptors and histograms for num_tasks tasks; the spill buffer
  // is sized on first use from the call's lengths.
  StageCounts c;
  c.nl = (n > (int64_t)f->small_cap) ? num_tasks : 0;
  c.ns = (n <= (int64_t)f->small_cap) ? num_tasks : 0;
  rc = w->pipe.reserve(c, 0, (hipStream_t)stream);
  if (rc) return rc;
  *ws = w.release();
  return 0;
}

int radik_ws_destroy(radik_ws* ws) {
  clear_error();
  if (!ws) return 0;
  int cur = 0;
  const int dev = ws->device;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  // Only the stream of the last call can still use the buffers (the side
  // stream's work joins it); other streams on the device are not waited for.
  if (ws->used) (void)hipStreamSynchronize(ws->last_stream);
  delete ws;
  if (cur != dev) (void)hipSetDevice(cur);
  return 0;
}

static int run_tasks(Pipeline& pipe, std::vector<TaskSpec>& specs, const LaunchOpts& o, hipStream_t s,
                     PhaseTimer* tm) {
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  rc = stage_tasks(pipe, specs.data(), (int64_t)specs.size(), *f, o, s);
  if (rc) return rc;
  rc = upload(pipe, s);
  if (rc) return rc;
  return enqueue(pipe, *f, o, s, tm);
}

static int ws_select(radik_ws* ws, void* val_out, void* idx_out, int idx_bytes, const void* val_in,
                     const void* idx_in, const int64_t* task_lens, int64_t num_tasks, int64_t k, uint32_t flags,
                     int vload_bytes, void* stream) {
  if (!ws) return set_error(RADIK_EINVAL, "workspace is NULL");
  if (k < 0 || k > INT32_MAX) return set_error(RADIK_EINVAL, "k = %lld out of range", (long long)k);
  if (num_tasks < 0) return set_error(RADIK_EINVAL, "num_tasks < 0");
  if (num_tasks > 0 && !task_lens) return set_error(RADIK_EINVAL, "task_lens is NULL");
  if (idx_bytes != 4 && idx_bytes != 8) return set_error(RADIK_EINVAL, "idx_bytes must be 4 or 8");
  if (flags & ~kFlagMask) return set_error(RADIK_EINVAL, "unknown flag bits 0x%x", flags & ~kFlagMask);
  int rc = check_key_flags(flags);
  if (rc) return rc;
  int64_t total = 0, maxlen = 0;
  rc = check_lens(task_lens, num_tasks, idx_bytes, &total, &maxlen);
  if (rc) return rc;
  if (maxlen > ws->n)
    return set_error(RADIK_EWORKSPACE, "workspace stride n = %lld < max(task_lens) = %lld", (long long)ws->n,
                     (long long)maxlen);
  if (num_tasks > ws->num_tasks)
    return set_error(RADIK_EWORKSPACE, "workspace holds %lld tasks, call has %lld", (long long)ws->num_tasks,
                     (long long)num_tasks);
  if (ws->counter_bytes == 4 && maxlen > INT32_MAX)
    return set_error(RADIK_EWORKSPACE, "Int32 workspace cannot count %lld elements", (long long)maxlen);
  if (k == 0 || num_tasks == 0) return 0;
  if (!val_out || !idx_out) return set_error(RADIK_EINVAL, "val_out/idx_out is NULL");
  if (total > 0 && !val_in) return set_error(RADIK_EINVAL, "val_in is NULL");
  if ((flags & RADIK_WITHIDXIN) && total > 0 && !idx_in)
    return set_error(RADIK_EINVAL, "WITHIDXIN set but idx_in is NULL");
  int cur = 0;
  RADIK_HIP(hipGetDevice(&cur));
  if (cur != ws->device) return set_error(RADIK_EINVAL, "workspace belongs to device %d, current is %d",
                                          ws->device, cur);
  hipStream_t s = (hipStream_t)stream;
  const uint32_t tflags = (flags & 0x0Fu) | (idx_bytes == 8 ? F_IDX64 : 0u);
  const int eb = key_bytes(key_type_of(flags));
  ws->specs.resize((size_t)num_tasks);
  int64_t off = 0;
  for (int64_t t = 0; t < num_tasks; ++t) {
    TaskSpec& sp = ws->specs[(size_t)t];
    sp.in = elem_ptr(val_in, off, eb);
    sp.idx_in = (flags & RADIK_WITHIDXIN) ? (const void*)((const char*)idx_in + off * idx_bytes) : nullptr;
    sp.val_out = elem_ptr(val_out, t * k, eb);
    sp.idx_out = (void*)((char*)idx_out + t * k * idx_bytes);
    sp.len = task_lens[t];
    sp.k = k;
    sp.flags = tflags;
    off += task_lens[t];
  }
  ws->last_stream = s;
  ws->used = true;
  if (key_type_of(flags) == KEY_F64) {
    if (k > RADIK_F64_MAX_K)
      return set_error(RADIK_EINVAL, "Float64 keys take k <= %d, got %lld", RADIK_F64_MAX_K, (long long)k);
    DevFacts* f = nullptr;
    if ((rc = dev_facts(&f))) return rc;
    if ((rc = f64_stage(&ws->f64, ws->specs.data(), num_tasks, s))) return rc;
    rc = f64_enqueue(ws->f64, s);
  } else {
    rc = run_tasks(ws->pipe, ws->specs, opts_for(flags, vload_bytes), s, nullptr);
  }
  if (rc) return rc;
  RADIK_HIP(hipStreamSynchronize(s));
  return 0;
}

int radik_topk_radix_select(radik_ws* ws, float* val_out, void* idx_out, int idx_bytes, const float* val_in,
                            const void* idx_in, const int64_t* task_lens, int64_t num_tasks, int64_t k,
                            uint32_t flags, int vload_bytes, void* stream) {
  clear_error();
  return ws_select(ws, val_out, idx_out, idx_bytes, val_in, idx_in, task_lens, num_tasks, k, flags, vload_bytes,
                   stream);
}

static int topk_impl(int key_type, void* val_out, void* idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
                     const void* val_in, const void* indices, int64_t n_indices, const int64_t* task_lens,
                     int64_t num_tasks, int64_t k, int largest, int rev, void* stream) {
  if (key_type < 0 || (key_type >= NUM_KEY_TYPES && key_type != KEY_F64))
    return set_error(RADIK_EINVAL, "unknown key type %d", key_type);
  if (num_tasks < 0 || (num_tasks > 0 && !task_lens)) return set_error(RADIK_EINVAL, "bad task_lens");
  if (out_rows != k || out_cols != num_tasks)
    return set_error(RADIK_ESIZE, "val_out must have size (%lld, %lld), got (%lld, %lld)", (long long)k,
                     (long long)num_tasks, (long long)out_rows, (long long)out_cols);
  int64_t total = 0, maxlen = 0;
  int rc = check_lens(task_lens, num_tasks, idx_bytes == 8 ? 8 : 4, &total, &maxlen);
  if (rc) return rc;
  if (indices && n_indices > 0 && n_indices != total)
    return set_error(RADIK_EINDICES, "indices length must equal data length. Got %lld vs %lld",
                     (long long)n_indices, (long long)total);
  uint32_t flags = RADIK_WITHPACKING | (largest ? RADIK_LARGEST : 0u) | (rev ? 0u : RADIK_ASCEND) |
                   ((uint32_t)key_type << RADIK_KEY_SHIFT);
  if (indices && n_indices > 0) flags |= RADIK_WITHIDXIN;
  // The reference allocates a workspace per call (algorithm.jl:503).  Here
  // each device keeps one cached workspace for these calls (grown as needed,
  // held until the process ends); a call that finds it in use by another
  // thread makes a temporary one.
  int dev = 0;
  RADIK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(RADIK_EINVAL, "device ordinal %d out of range", dev);
  TopkCache& tc = g_topk_cache[dev];
  std::unique_lock<std::mutex> lk(tc.mu, std::try_to_lock);
  radik_ws* ws = nullptr;
  bool temp = !lk.owns_lock();
  if (!temp && !tc.ws) {
    rc = radik_ws_create(&tc.ws, maxlen, num_tasks, 4, stream);
    if (rc) return rc;
  }
  if (temp) {
    rc = radik_ws_create(&ws, maxlen, num_tasks, 4, stream);
    if (rc) return rc;
  } else {
    ws = tc.ws;
    if (ws->n < maxlen) ws->n = maxlen;
    if (ws->num_tasks < num_tasks) ws->num_tasks = num_tasks;
  }
  rc = ws_select(ws, val_out, idx_out, idx_bytes, val_in, indices, task_lens, num_tasks, k, flags, 32, stream);
  std::string keep = rc ? std::string(radik_last_error()) : std::string();
  if (temp) radik_ws_destroy(ws);
  if (rc) return set_error(rc, "%s", keep.c_str());
  return 0;
}

int radik_topk(float* val_out, void* idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
               const float* val_in, const void* indices, int64_t n_indices, const int64_t* task_lens,
               int64_t num_tasks, int64_t k, int largest, int rev, void* stream) {
  clear_error();
  return topk_impl(KEY_F32, val_out, idx_out, idx_bytes, out_rows, out_cols, val_in, indices, n_indices, task_lens,
                   num_tasks, k, largest, rev, stream);
}

int radik_topk_typed(int key_type, void* val_out, void* idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
                     const void* val_in, const void* indices, int64_t n_indices, const int64_t* task_lens,
                     int64_t num_tasks, int64_t k, int largest, int rev, void* stream) {
  clear_error();
  return topk_impl(key_type, val_out, idx_out, idx_bytes, out_rows, out_cols, val_in, indices, n_indices,
                   task_lens, num_tasks, k, largest, rev, stream);
}

int radik_topk_noalloc(float* val_out, void* idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
                       const float* val_in, const void* indices, int64_t n_indices, const int64_t* task_lens,
                       int64_t num_tasks, int64_t k, int largest, int rev, void* stream) {
  clear_error();
  if (num_tasks < 0 || (num_tasks > 0 && !task_lens)) return set_error(RADIK_EINVAL, "bad task_lens");
  if (out_rows != k || out_cols != num_tasks)
    return set_error(RADIK_ESIZE, "val_out must have size (%lld, %lld), got (%lld, %lld)", (long long)k,
                     (long long)num_tasks, (long long)out_rows, (long long)out_cols);
  if (k < 0 || k > INT32_MAX) return set_error(RADIK_EINVAL, "k = %lld out of range", (long long)k);
  if (idx_bytes != 4 && idx_bytes != 8) return set_error(RADIK_EINVAL, "idx_bytes must be 4 or 8");
  int64_t total = 0, maxlen = 0;
  int rc = check_lens(task_lens, num_tasks, idx_bytes, &total, &maxlen);
  if (rc) return rc;
  if (indices && n_indices > 0 && n_indices != total)
    return set_error(RADIK_EINDICES, "indices length must equal data length. Got %lld vs %lld",
                     (long long)n_indices, (long long)total);
  if (num_tasks > RADIK_NOALLOC_MAX_TASKS)
    return set_error(RADIK_ELIMIT, "radik_topk_noalloc takes at most %d tasks, got %lld", RADIK_NOALLOC_MAX_TASKS,
                     (long long)num_tasks);
  if (total > RADIK_NOALLOC_MAX_ELEMS)
    return set_error(RADIK_ELIMIT, "radik_topk_noalloc takes at most %d elements in all, got %lld",
                     RADIK_NOALLOC_MAX_ELEMS, (long long)total);
  int64_t nlong = 0;
  for (int64_t t = 0; t < num_tasks; ++t) nlong += task_lens[t] > 2048;
  if (nlong > RADIK_NOALLOC_MAX_LONG)
    return set_error(RADIK_ELIMIT, "radik_topk_noalloc takes at most %d tasks longer than 2048 elements, got %lld",
                     RADIK_NOALLOC_MAX_LONG, (long long)nlong);
  if (k == 0 || num_tasks == 0) return 0;
  if (!val_out || !idx_out) return set_error(RADIK_EINVAL, "val_out/idx_out is NULL");
  if (total > 0 && !val_in) return set_error(RADIK_EINVAL, "val_in is NULL");
  const bool withidx = indices && n_indices > 0;
  const uint32_t tflags = RADIK_WITHPACKING | (largest ? RADIK_LARGEST : 0u) | (rev ? 0u : RADIK_ASCEND) |
                          (withidx ? RADIK_WITHIDXIN : 0u);
  return noalloc_run(val_out, idx_out, idx_bytes, val_in, withidx ? indices : nullptr, task_lens, num_tasks, k,
                     (tflags & 0x0Fu) | (idx_bytes == 8 ? F_IDX64 : 0u), (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Batched plans
// ---------------------------------------------------------------------------
int radik_plan_create(radik_plan** plan, const int64_t* task_lens, const int64_t* k_per_task, int64_t num_tasks,
                      uint32_t flags, int idx_bytes, void* stream) {
  clear_error();
  if (!plan) return set_error(RADIK_EINVAL, "plan is NULL");
  *plan = nullptr;
  if (num_tasks < 0 || (num_tasks > 0 && (!task_lens || !k_per_task)))
    return set_error(RADIK_EINVAL, "bad task_lens / k_per_task");
  if (idx_bytes != 4 && idx_bytes != 8) return set_error(RADIK_EINVAL, "idx_bytes must be 4 or 8");
  if (flags & ~kFlagMask) return set_error(RADIK_EINVAL, "unknown flag bits 0x%x", flags & ~kFlagMask);
  int rc = check_key_flags(flags);
  if (rc) return rc;
  if (key_type_of(flags) == KEY_F64)
    return set_error(RADIK_EINVAL, "Float64 keys run through radik_topk_typed and the workspace path, not plans");
  int64_t total = 0, maxlen = 0;
  rc = check_lens(task_lens, num_tasks, idx_bytes, &total, &maxlen);
  if (rc) return rc;
  std::unique_ptr<radik_plan> p(new (std::nothrow) radik_plan());
  if (!p) return set_error(RADIK_ENOMEM, "out of host memory");
  RADIK_HIP(hipGetDevice(&p->device));
  p->lens.assign(task_lens, task_lens + num_tasks);
  p->ks.assign(k_per_task, k_per_task + num_tasks);
  p->flags = flags;
  p->idx_bytes = idx_bytes;
  p->opts = opts_for(flags | RADIK_WITHPACKING, 32);
  p->in_off.resize((size_t)num_tasks);
  p->out_off.resize((size_t)num_tasks);
  int64_t io = 0, oo = 0;
  for (int64_t t = 0; t < num_tasks; ++t) {
    if (p->ks[(size_t)t] < 0 || p->ks[(size_t)t] > INT32_MAX)
      return set_error(RADIK_EINVAL, "k_per_task[%lld] out of range", (long long)t);
    p->in_off[(size_t)t] = io;
    p->out_off[(size_t)t] = oo;
    io += p->lens[(size_t)t];
    oo += p->ks[(size_t)t];
  }
  p->specs.resize((size_t)num_tasks);
  (void)stream;
  *plan = p.release();
  return 0;
}

int radik_plan_set_graph(radik_plan* p, int use_graph) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  p->use_graph = use_graph != 0;
  return 0;
}

int radik_plan_set_options(radik_plan* p, int vload_bytes, int blocks_per_cu, int small_cap, int refine_cap,
                           int predict) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  if (blocks_per_cu < 0 || blocks_per_cu > 64 || small_cap < 0 || refine_cap < 0)
    return set_error(RADIK_EINVAL, "bad plan options");
  if (vload_bytes > 0) p->opts = opts_for(p->flags | RADIK_WITHPACKING, vload_bytes);
  if (blocks_per_cu > 0) p->opts.blocks_per_cu = blocks_per_cu;
  p->opts.small_cap = (uint32_t)small_cap;
  p->opts.refine_cap = (uint32_t)refine_cap;
  p->opts.predict = predict != 0;
  p->staged = false;  // re-stage (task order may change) and re-capture
  return 0;
}

int radik_plan_set_option(radik_plan* p, const char* name, int64_t value) {
  clear_error();
  if (!p || !name) return set_error(RADIK_EINVAL, "plan or name is NULL");
  const std::string n(name);
  if (n == "medium") {
    p->opts.medium = value != 0;

  } else if (n == "fork_late") {
    if (value < -1 || value > 2) return set_error(RADIK_EINVAL, "fork_late must be -1, 0, 1 or 2");
    p->opts.fork_late = (int)value;
  } else if (n == "longest_first") {
    if (value < 0 || value > 1) return set_error(RADIK_EINVAL, "longest_first must be 0 or 1");
    p->opts.longest_first = value != 0;
  } else if (n == "join_early") {
    if (value < -1 || value > 1) return set_error(RADIK_EINVAL, "join_early must be -1, 0 or 1");
    p->opts.join_early = (int)value;
  } else if (n == "sample_min_len") {
    if (value < 4096 || value > ((int64_t)1 << 30)) return set_error(RADIK_EINVAL, "sample_min_len must be in 4096..2^30");
    p->opts.sample_min_len = value;
  } else if (n == "medium_per_cu") {
    if (value < 1 || value > 16) return set_error(RADIK_EINVAL, "medium_per_cu must be in 1..16");
    p->opts.med_per_cu = (int)value;
  } else if (n == "predict") {
    p->opts.predict = value != 0;
  } else if (n == "blocks_per_cu") {
    if (value < 1 || value > 64) return set_error(RADIK_EINVAL, "blocks_per_cu must be in 1..64");
    p->opts.blocks_per_cu = (int)value;
  } else if (n == "small_cap") {
    if (value < 0 || value > INT32_MAX) return set_error(RADIK_EINVAL, "bad small_cap");
    p->opts.small_cap = (uint32_t)value;
  } else if (n == "refine_cap") {
    if (value < 0 || value > INT32_MAX) return set_error(RADIK_EINVAL, "bad refine_cap");
    p->opts.refine_cap = (uint32_t)value;
  } else if (n == "vload_bytes") {
    if (value != 16 && value != 32 && value != 64) return set_error(RADIK_EINVAL, "vload_bytes must be 16, 32 or 64");
    const LaunchOpts keep = p->opts;
    p->opts = opts_for(p->flags | RADIK_WITHPACKING, (int)value);
    p->opts.blocks_per_cu = keep.blocks_per_cu;
    p->opts.small_cap = keep.small_cap;
    p->opts.refine_cap = keep.refine_cap;
    p->opts.predict = keep.predict;
    p->opts.medium = keep.medium;
    p->opts.fork_late = keep.fork_late;
    p->opts.join_early = keep.join_early;
    p->opts.longest_first = keep.longest_first;
    p->opts.med_per_cu = keep.med_per_cu;
    p->opts.sample_min_len = keep.sample_min_len;
  } else {
    return set_error(RADIK_EINVAL, "unknown plan option '%s'", name);
  }
  p->staged = false;  // re-stage (routing and task order may change) and re-capture
  return 0;
}

int radik_plan_set_timing(radik_plan* p, int enable) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  if (enable && !p->timer.ev[0])
    for (auto& e : p->timer.ev) RADIK_HIP(hipEventCreate(&e));
  p->timer.on = enable != 0;
  for (double& m : p->ms) m = 0;
  p->timed = 0;
  return 0;
}

int radik_plan_timing(radik_plan* p, double* ms8, int64_t* count) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  if (ms8)
    for (int i = 0; i < NUM_PHASES; ++i) ms8[i] = p->ms[i];
  if (count) *count = p->timed;
  return 0;
}

int radik_plan_execute(radik_plan* p, void* val_out, void* idx_out, const void* val_in, const void* idx_in,
                       void* stream) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  const int64_t T = (int64_t)p->lens.size();
  if (T == 0) return 0;
  if (!val_out || !idx_out || !val_in) return set_error(RADIK_EINVAL, "NULL buffer");
  if ((p->flags & RADIK_WITHIDXIN) && !idx_in) return set_error(RADIK_EINVAL, "WITHIDXIN set but idx_in is NULL");
  hipStream_t s = (hipStream_t)stream;
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  const bool same = p->staged && p->last_in == val_in && p->last_idx_in == idx_in && p->last_val_out == val_out &&
                    p->last_idx_out == idx_out;
  if (!same) {
    const uint32_t tflags = (p->flags & 0x0Fu) | (p->idx_bytes == 8 ? F_IDX64 : 0u);
    const int eb = key_bytes(key_type_of(p->flags));
    for (int64_t t = 0; t < T; ++t) {
      TaskSpec& sp = p->specs[(size_t)t];
      const int64_t io = p->in_off[(size_t)t], oo = p->out_off[(size_t)t];
      sp.in = elem_ptr(val_in, io, eb);
      sp.idx_in = (p->flags & RADIK_WITHIDXIN) ? (const void*)((const char*)idx_in + io * p->idx_bytes) : nullptr;
      sp.val_out = elem_ptr(val_out, oo, eb);
      sp.idx_out = (void*)((char*)idx_out + oo * p->idx_bytes);
      sp.len = p->lens[(size_t)t];
      sp.k = p->ks[(size_t)t];
      sp.flags = tflags;
    }
    rc = stage_tasks(p->pipe, p->specs.data(), T, *f, p->opts, s);
    if (rc) return rc;
    rc = upload(p->pipe, s);
    if (rc) return rc;
    p->last_in = val_in;
    p->last_idx_in = idx_in;
    p->last_val_out = val_out;
    p->last_idx_out = idx_out;
    p->staged = true;
  }
  if (p->timer.on) {
    rc = enqueue(p->pipe, *f, p->opts, s, &p->timer);
    if (rc) return rc;
    RADIK_HIP(hipEventSynchronize(p->timer.ev[NUM_PHASES]));
    for (int i = 0; i < NUM_PHASES; ++i) {
      float ms = 0;
      RADIK_HIP(hipEventElapsedTime(&ms, p->timer.ev[i], p->timer.ev[i + 1]));
      p->ms[i] += ms;
    }
    p->timed++;
    return 0;
  }
  if (!p->use_graph) return enqueue(p->pipe, *f, p->opts, s, nullptr);
  // Device addresses in the graph (descriptor arrays, histogram, spill) are
  // fixed once staged; a buffer reallocation or another stream re-captures.
  if (!p->exec || p->graph_stream != s || !same) {
    if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
    if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
    hipStream_t cs = s;
    bool own = false;
    if (!cs) {  // the legacy default stream cannot be captured
      RADIK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      own = true;
    }
    if (own) RADIK_HIP(hipStreamSynchronize(nullptr));
    RADIK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    LaunchOpts go = p->opts;
    go.graph = true;
    rc = enqueue(p->pipe, *f, go, cs, nullptr);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(cs, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); if (own) (void)hipStreamDestroy(cs); return rc; }
    if (ce != hipSuccess) {
      if (own) (void)hipStreamDestroy(cs);
      return set_error(RADIK_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
    }
    p->graph = g;
    hipError_t ie = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(cs);
    if (ie != hipSuccess) return set_error(RADIK_EHIP, "graph instantiate failed: %s", hipGetErrorString(ie));
    p->graph_stream = s;
  }
  RADIK_HIP(hipGraphLaunch(p->exec, s));
  return 0;
}

int radik_plan_sample(radik_plan* p, const void* val_out, const void* idx_out, float temperature, float top_p,
                      const float* uniforms, int64_t* tokens, int32_t* ranks, void* stream) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "plan is NULL");
  if (!(p->flags & RADIK_LARGEST) || (p->flags & RADIK_ASCEND))
    return set_error(RADIK_EINVAL, "sampling needs a plan with largest = true and rev = true (descending results)");
  const int64_t T = (int64_t)p->lens.size();
  if (T == 0) return 0;
  if (!val_out || !idx_out || !tokens || !ranks) return set_error(RADIK_EINVAL, "NULL buffer");
  if (temperature > 0.0f && !uniforms) return set_error(RADIK_EINVAL, "uniforms is NULL");
  if (!(top_p >= 0.0f)) return set_error(RADIK_EINVAL, "top_p must be >= 0");
  if (!p->d_soff) {
    std::vector<int32_t> m((size_t)T);
    for (int64_t t = 0; t < T; ++t) m[(size_t)t] = (int32_t)std::min(p->lens[(size_t)t], p->ks[(size_t)t]);
    if (hipMalloc((void**)&p->d_soff, (size_t)T * sizeof(int64_t)) != hipSuccess ||
        hipMalloc((void**)&p->d_sm, (size_t)T * sizeof(int32_t)) != hipSuccess)
      return set_error(RADIK_ENOMEM, "allocation of the sampling tables failed");
    RADIK_HIP(hipMemcpy(p->d_soff, p->out_off.data(), (size_t)T * sizeof(int64_t), hipMemcpyHostToDevice));
    RADIK_HIP(hipMemcpy(p->d_sm, m.data(), (size_t)T * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  return sample_topp(val_out, idx_out, p->idx_bytes, p->d_soff, p->d_sm, T, key_type_of(p->flags), temperature,
                     top_p, uniforms, tokens, ranks, (hipStream_t)stream);
}

int radik_plan_info(radik_plan* p, int64_t* info, int n) {
  clear_error();
  if (!p || (n > 0 && !info)) return set_error(RADIK_EINVAL, "plan or info is NULL");
  if (!p->staged) return set_error(RADIK_EINVAL, "the plan has not been executed yet");
  const Pipeline& q = p->pipe;
  const int64_t v[RADIK_PLAN_INFO_LEN] = {q.n_large, q.total, q.n_med, q.total_med, q.n_small, q.total_small,
                                          q.n_chunks, q.total_groups, q.has_big ? 1 : 0};
  for (int i = 0; i < n; ++i) info[i] = i < RADIK_PLAN_INFO_LEN ? v[i] : 0;
  return 0;
}

int radik_plan_destroy(radik_plan* p) {
  clear_error();
  if (!p) return 0;
  (void)hipDeviceSynchronize();
  delete p;
  return 0;
}

// ---------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------
int radik_diag_scan(const uint32_t* in, uint32_t* out_wave, uint32_t* out_block, int64_t n, void* stream) {
  clear_error();
  if (n < 0 || (n > 0 && (!in || !out_wave || !out_block))) return set_error(RADIK_EINVAL, "bad arguments");
  int rc = diag_scan(in, out_wave, out_block, n, (hipStream_t)stream);
  if (rc) return rc;
  RADIK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int radik_diag_pass1(const float* val_in, const int64_t* task_lens, int64_t num_tasks, int64_t k, uint32_t flags,
                     uint32_t* hist_out, uint32_t* sel_out, void* stream) {
  clear_error();
  if (num_tasks <= 0 || !task_lens || !val_in || !hist_out || !sel_out) return set_error(RADIK_EINVAL, "bad arguments");
  if (flags & RADIK_KEY_MASK) return set_error(RADIK_EINVAL, "radik_diag_pass1 takes Float32 keys");
  if (k < 0 || k > INT32_MAX) return set_error(RADIK_EINVAL, "bad k");
  int64_t total = 0, maxlen = 0;
  int rc = check_lens(task_lens, num_tasks, 8, &total, &maxlen);
  if (rc) return rc;
  DevFacts* f = nullptr;
  rc = dev_facts(&f);
  if (rc) return rc;
  // Every task goes through the streaming pass here, short or not.
  DevFacts all = *f;
  all.small_cap = 0;
  Pipeline pipe;
  std::vector<TaskSpec> specs((size_t)num_tasks);
  int64_t off = 0;
  for (int64_t t = 0; t < num_tasks; ++t) {
    TaskSpec& sp = specs[(size_t)t];
    sp.in = val_in + off;
    sp.idx_in = nullptr;
    sp.val_out = nullptr;
    sp.idx_out = nullptr;
    sp.len = task_lens[t];
    sp.k = k;
    sp.flags = flags & 0x0Fu;
    off += task_lens[t];
  }
  hipStream_t s = (hipStream_t)stream;
  LaunchOpts o;
  o.predict = false;
  o.medium = false;
  o.longest_first = false;  // (the outputs are per staged task, in the caller's order)
  rc = stage_tasks(pipe, specs.data(), num_tasks, all, o, s);
  if (rc) return rc;
  rc = upload(pipe, s);
  if (rc) return rc;
  rc = diag_pass1(pipe, all, hist_out, sel_out, s);
  if (rc) return rc;
  RADIK_HIP(hipStreamSynchronize(s));
  return 0;
}

int radik_diag_plan_chunks(int64_t total, int64_t blocks, int64_t spread, int guided, int64_t gran,
                           int64_t min_chunk, int64_t* starts, int64_t cap, int64_t* nchunks) {
  clear_error();
  if (!nchunks || total < 0 || blocks < 1 || spread < 1 || gran < 1 || min_chunk < 1)
    return set_error(RADIK_EINVAL, "bad arguments");
  *nchunks = diag_plan_chunks(total, blocks, spread, guided, gran, min_chunk, starts, cap);
  if (starts && cap < *nchunks + 1) return set_error(RADIK_ESIZE, "starts needs %lld entries", (long long)(*nchunks + 1));
  return 0;
}

int radik_diag_route(const int64_t* task_lens, const int64_t* k_per_task, int64_t num_tasks, int small_cap,
                     int predict, int medium, int32_t* cls, int64_t* groups) {
  clear_error();
  if (num_tasks < 0 || (num_tasks > 0 && (!task_lens || !k_per_task || !cls || !groups)) || small_cap < 0)
    return set_error(RADIK_EINVAL, "bad arguments");
  diag_route(task_lens, k_per_task, num_tasks, (uint32_t)small_cap, predict, medium, cls, groups);
  return 0;
}

int radik_diag_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  clear_error();
  if (bytes < 0 || (bytes > 0 && (!dst || !src)) || (bytes % 16) != 0)
    return set_error(RADIK_EINVAL, "bytes must be a multiple of 16 with non-NULL buffers");
  int rc = diag_copy(dst, src, bytes, (hipStream_t)stream);
  if (rc) return rc;
  RADIK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

// ---------------------------------------------------------------------------
// Device-memory helpers
// ---------------------------------------------------------------------------
int radik_device_count(int* n) {
  clear_error();
  if (!n) return set_error(RADIK_EINVAL, "NULL");
  RADIK_HIP(hipGetDeviceCount(n));
  return 0;
}

int radik_set_device(int dev) {
  clear_error();
  RADIK_HIP(hipSetDevice(dev));
  return 0;
}

int radik_device_info(int* cus, int64_t* lds_per_block, int64_t* hbm_bytes, char* name, int name_len) {
  clear_error();
  int dev = 0;
  RADIK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  RADIK_HIP(hipGetDeviceProperties(&pr, dev));
  if (cus) *cus = pr.multiProcessorCount;
  if (lds_per_block) *lds_per_block = (int64_t)pr.sharedMemPerBlock;
  if (hbm_bytes) *hbm_bytes = (int64_t)pr.totalGlobalMem;
  if (name && name_len > 0) {
    strncpy(name, pr.gcnArchName, (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  return 0;
}

int radik_lds_caps(uint32_t* small_cap, uint32_t* refine_cap, uint32_t* sort_cap) {
  clear_error();
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  if (small_cap) *small_cap = f->small_cap;
  if (refine_cap) *refine_cap = f->refine_cap;
  if (sort_cap) *sort_cap = f->sort_cap;
  return 0;
}

int radik_malloc(void** p, size_t bytes) {
  clear_error();
  if (!p) return set_error(RADIK_EINVAL, "NULL");
  *p = nullptr;
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return set_error(RADIK_ENOMEM, "hipMalloc(%zu) failed", bytes);
  return 0;
}

int radik_free(void* p) {
  clear_error();
  if (p) RADIK_HIP(hipFree(p));
  return 0;
}

int radik_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  clear_error();
  if (bytes) RADIK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

int radik_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  clear_error();
  if (bytes) RADIK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int radik_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  clear_error();
  if (bytes) RADIK_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return 0;
}

int radik_memset(void* p, int v, size_t bytes) {
  clear_error();
  if (bytes) RADIK_HIP(hipMemset(p, v, bytes));
  return 0;
}

int radik_device_sync(void) {
  clear_error();
  RADIK_HIP(hipDeviceSynchronize());
  return 0;
}

int radik_stream_create(void** s) {
  clear_error();
  if (!s) return set_error(RADIK_EINVAL, "NULL");
  hipStream_t st = nullptr;
  RADIK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = st;
  return 0;
}

int radik_stream_destroy(void* s) {
  clear_error();
  if (s) RADIK_HIP(hipStreamDestroy((hipStream_t)s));
  return 0;
}

int radik_stream_sync(void* s) {
  clear_error();
  RADIK_HIP(hipStreamSynchronize((hipStream_t)s));
  return 0;
}

int radik_event_create(void** e) {
  clear_error();
  if (!e) return set_error(RADIK_EINVAL, "NULL");
  hipEvent_t ev = nullptr;
  RADIK_HIP(hipEventCreate(&ev));
  *e = ev;
  return 0;
}

int radik_event_destroy(void* e) {
  clear_error();
  if (e) RADIK_HIP(hipEventDestroy((hipEvent_t)e));
  return 0;
}

int radik_event_record(void* e, void* s) {
  clear_error();
  RADIK_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
  return 0;
}

int radik_event_sync(void* e) {
  clear_error();
  RADIK_HIP(hipEventSynchronize((hipEvent_t)e));
  return 0;
}

int radik_event_elapsed(void* a, void* b, float* ms) {
  clear_error();
  if (!ms) return set_error(RADIK_EINVAL, "NULL");
  RADIK_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return 0;
}

}  // extern "C"
