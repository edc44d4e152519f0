// radik_host.h -- host-side pieces shared by the C-ABI translation units:
// error reporting, device facts, the launch pipeline and its buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "radik_device.h"

namespace radik {

// Thread-local error slot behind radik_last_error().
int set_error(int code, const char* fmt, ...);
void clear_error();

#define RADIK_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::radik::set_error(5, "%s failed: %s", #call, hipGetErrorString(e_));  \
  } while (0)

// Facts of the current device, queried once per device.
struct DevFacts {
  int device = -1;
  int cus = 256;
  size_t lds_per_block = 65536;
  uint32_t small_cap = 2048;   // default: longest task routed to k_small
  uint32_t refine_cap = 4096;  // default: candidates kept in LDS by k_finish
  uint32_t small_cap_max = 2048, refine_cap_max = 4096;  // what the LDS budget allows
  uint32_t sort_cap = 4096;    // results sorted in LDS by k_sort
  int hist_blocks_per_cu[NUM_KEY_TYPES] = {4, 4, 4, 4};  // resident pass-A blocks per CU, per key type
};
int dev_facts(DevFacts** out);

// Tasks up to BIG_LEN elements without a usable spill are re-read by one
// k_finish block; longer ones by pass B over several blocks (k_compact).
constexpr int64_t BIG_LEN = (int64_t)1 << 20;
// Static LDS of k_finish beside its dynamic arrays (upper bound, bytes).
constexpr size_t FINISH_STATIC_LDS = 2048;
// Medium tasks (longer than the short cap, at most MED_LEN elements, not
// sampled, m <= MED_CAP) run whole in one k_medium block beside the passes.
constexpr int64_t MED_LEN = (int64_t)1 << 16;
constexpr uint32_t MED_CAP = 2048;
// Medium-task blocks resident per CU at most (by their LDS request).
constexpr int MED_MAX_PER_CU = 5;

// Launch-level options.
struct LaunchOpts {
  int unr = 2;              // float4 loads in flight per thread (vload bytes / 16)
  int blocks_per_cu = 8;    // streaming grid: blocks per CU
  uint32_t small_cap = 0;   // longest task for k_small (0: device default)
  uint32_t refine_cap = 0;  // candidates held in LDS by k_finish (0: device default)
  bool predict = true;      // sample long tasks and spill in pass A
  bool medium = true;       // medium tasks to k_medium (else the streaming passes)
  int fork_late = -1;       // short/medium tasks start with the launch (0), after the prediction (1)
                            // or after the sampling (2); -1: 2
  bool longest_first = true;  // tasks staged longest first (else in the caller's order)
  int join_early = -1;      // the launch's stream waits for the short/medium tasks before k_finish (1)
                            // or at the end (0); -1: 1 when the long tasks hold at least as many
                            // elements as the short and medium ones
  bool graph = false;       // the sequence is being captured into a graph
  int med_per_cu = MED_MAX_PER_CU;  // k_medium blocks that may share a CU (LDS reserved per block)
  int64_t sample_min_len = (int64_t)1 << 15;  // shortest task that is sampled and predicted
  int key_type = KEY_F32;   // element type of every task of the launch (KEY_*)
};
inline uint32_t eff_small_cap(const DevFacts& f, const LaunchOpts& o) {
  return o.small_cap ? (o.small_cap < f.small_cap_max ? o.small_cap : f.small_cap_max) : f.small_cap;
}
inline uint32_t eff_refine_cap(const DevFacts& f, const LaunchOpts& o) {
  return o.refine_cap ? (o.refine_cap < f.refine_cap_max ? o.refine_cap : f.refine_cap_max) : f.refine_cap;
}

// Counts that size a launch's descriptor arrays (see Pipeline).
struct StageCounts {
  int64_t nl = 0, ns = 0, nm = 0;  // long, short, medium tasks
  int64_t nch = 0;                 // pass A's chunks
  int64_t nbmap = 0;               // block-map entries (chunks + sampling blocks)
  int64_t nsplit = 0;              // k_split blocks
  int64_t nrt = 0;                 // rank-tail runs (one run-sort block each)
};

// Device + host buffers of one pipeline instance.  Every array the host
// builds for a launch (descriptors, prefix sums, chunk table, block maps)
// lives in one arena, copied to the device with one copy; the control
// blocks follow it on the device.  Histograms (kept all-zero between
// launches: the consumer clears what it reads) and spill regions are
// separate.  Grow-only.  A fixed pipeline (bind_fixed) lives in storage the
// caller provides and never allocates: a launch that does not fit fails.
struct Pipeline {
  // host copies (pinned, or the caller's storage when fixed)
  TaskDesc* h_large = nullptr;
  int64_t* h_bases = nullptr;   // element prefix sums of the long tasks
  int64_t* h_sbases = nullptr;  // sample-group prefix sums of the long tasks
  int64_t* h_cbases = nullptr;  // sorted-run prefix sums (multi-block sort)
  int64_t* h_mbases = nullptr;  // result-count prefix sums (multi-block sort)
  uint32_t* h_bmap = nullptr;   // first task of each chunk of the streaming launches
  int64_t* h_cstart = nullptr;  // pass A's chunk table: nch + 1 starts in the element space
  uint2* h_xmap = nullptr;      // k_split: task / block in task / blocks of task, per block
  uint2* h_rmap = nullptr;      // rank tail: task / run in task / runs of task, per run
  TaskDesc* h_small = nullptr;
  TaskDesc* h_med = nullptr;
  // device
  TaskDesc* d_large = nullptr;
  int64_t* d_bases = nullptr;
  int64_t* d_sbases = nullptr;
  int64_t* d_cbases = nullptr;
  int64_t* d_mbases = nullptr;
  uint32_t* d_bmap = nullptr;
  int64_t* d_cstart = nullptr;
  uint2* d_xmap = nullptr;
  uint2* d_rmap = nullptr;
  TaskDesc* d_small = nullptr;
  TaskDesc* d_med = nullptr;
  TaskCtl* d_ctl = nullptr;
  uint32_t* d_hist = nullptr;   // exact first-digit histograms, zero between launches
  uint32_t* d_shist = nullptr;  // sample histograms, zero between launches
  uint64_t* d_cand = nullptr;   // spill regions
  char* h_arena = nullptr;
  char* d_arena = nullptr;
  size_t cap_arena = 0;     // bytes of each arena
  size_t upload_bytes = 0;  // prefix of the arena the launch copies to the device
  size_t cap_hist = 0, cap_cand = 0;
  bool fixed = false;
  int64_t n_split = 0;  // k_split blocks (tasks longer than BIG_LEN)
  int64_t n_rt = 0;     // rank-tail runs (tasks longer than BIG_LEN, without WITHSCALE)
  uint32_t rt_lds = 0;  // LDS entries of a rank block: the longest list given, up to RT_LDS
  // Streaming launch shapes fixed at staging (passes A and B share one chunk
  // table, one chunk per block; sampling has equal chunks): the number of
  // chunks (= pass A's grid), the sampling grid, offsets of their maps in the
  // block map.
  int64_t n_chunks = 0;
  int grid_h = 0, grid_s = 0;
  int64_t off_h = 0, off_s = 0, n_bmap = 0;
  // current content
  int64_t n_large = 0, n_small = 0, total = 0, total_groups = 0;
  int64_t n_med = 0;
  int64_t total_med = 0, total_small = 0;  // elements of the medium and short tasks
  uint32_t max_m_med = 0;
  bool has_big = false;    // some long task exceeds BIG_LEN (pass B over several blocks)
  bool pos_bounds = false; // some predicted task may get a spill bound with a position (PRED_MAX_G)
  uint32_t max_m = 0;  // largest min(k, len) over the long tasks
  int64_t total_runs = 0, total_m = 0;  // multi-block sort sizes
  hipEvent_t upload_done = nullptr;  // pinned buffers reusable after this
  // k_small runs beside the long-task passes on a second stream (fork/join
  // events, also inside a graph capture); a fixed pipeline has none and runs
  // everything on the caller's stream.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  ~Pipeline();
  // tot_cand: spill-region entries (long and medium tasks)
  int reserve(const StageCounts& c, int64_t tot_cand, hipStream_t s);
  // Places the pipeline in caller storage: dev (device, 256-byte aligned)
  // holds histograms for max_long tasks, max_cand spill entries and the
  // arena; host holds the host arena.  Nothing is allocated afterwards.
  int bind_fixed(char* dev, size_t dev_bytes, char* host, size_t host_bytes, int64_t max_long, int64_t max_cand);
};

// Splits tasks into short ones (k_small) and long ones (streaming passes),
// fills the pinned descriptor arrays and their prefix sums.
struct TaskSpec {
  const void* in;  // elements of the launch's key type
  const void* idx_in;
  void* val_out;
  void* idx_out;
  int64_t len;
  int64_t k;
  uint32_t flags;
};
int stage_tasks(Pipeline& p, const TaskSpec* tasks, int64_t n, const DevFacts& f, const LaunchOpts& o,
                hipStream_t s);

// Address of element i of an array of eb-byte elements.
inline const void* elem_ptr(const void* base, int64_t i, int eb) {
  return base ? (const void*)((const char*)base + i * (int64_t)eb) : nullptr;
}
inline void* elem_ptr(void* base, int64_t i, int eb) {
  return base ? (void*)((char*)base + i * (int64_t)eb) : nullptr;
}

// Phase timing hooks (events recorded between kernels when non-null).
// Phases: sample, predict, hist, select, compact, finish, sort, small.
constexpr int NUM_PHASES = 8;
struct PhaseTimer {
  hipEvent_t ev[NUM_PHASES + 1] = {};
  bool on = false;
};

// Copies the staged descriptors to the device and enqueues every kernel.
int upload(Pipeline& p, hipStream_t s);
int enqueue(Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s, PhaseTimer* tm);

// The Float64 path (radik_f64.hip): its own grow-only buffers per workspace;
// staging copies the descriptors synchronously, enqueue launches the pass
// sequence on the stream.  f64_setup opts its sort into the LDS it needs.
struct F64Pipe;
int f64_stage(F64Pipe** p, const TaskSpec* specs, int64_t n, hipStream_t s);
int f64_enqueue(F64Pipe* p, hipStream_t s);
void f64_destroy(F64Pipe* p);
int f64_setup(size_t lds_budget);

}  // namespace radik
