// radik_pipeline.hip -- host orchestration of the kernel sequence.  The
// kernels are compiled in this translation unit so every launch resolves in
// one code object.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/radik.h"
#include "radik_host.h"

namespace radik {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

void clear_error() { g_err.clear(); }

const char* last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Device facts
// ---------------------------------------------------------------------------
static std::mutex g_facts_mu;
static DevFacts g_facts[64];

// The kernels of each key type (radik_launch.hip, one object per type).
#define RADIK_DECL_KT(NS)                                                                            \
  int launch_setup_##NS(size_t lds_budget, int* hist_occ);                                          \
  int enqueue_##NS(Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s, PhaseTimer* tm);
RADIK_DECL_KT(kf32)
RADIK_DECL_KT(kf16)
RADIK_DECL_KT(kbf16)
RADIK_DECL_KT(ki32)
#undef RADIK_DECL_KT

int dev_facts(DevFacts** out) {
  int dev = 0;
  RADIK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_error(5, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_facts_mu);
  DevFacts& f = g_facts[dev];
  if (f.device != dev) {
    int cus = 0, lds = 0;
    RADIK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    RADIK_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    f.cus = cus > 0 ? cus : 256;
    // LDS budget per block: the device's limit once every key type's
    // per-task kernels are opted into the larger allocation, else 64 KiB.
    size_t budget = (size_t)(lds > 0 ? lds : 65536);
    int occ[NUM_KEY_TYPES] = {4, 4, 4, 4};
    const bool ok = launch_setup_kf32(budget, &occ[KEY_F32]) == 0 && launch_setup_kf16(budget, &occ[KEY_F16]) == 0 &&
                    launch_setup_kbf16(budget, &occ[KEY_BF16]) == 0 && launch_setup_ki32(budget, &occ[KEY_I32]) == 0 &&
                    f64_setup(budget) == 0;
    if (!ok) {
      (void)hipGetLastError();
      budget = 65536;
    }
    f.lds_per_block = budget;
    for (int t = 0; t < NUM_KEY_TYPES; ++t) f.hist_blocks_per_cu[t] = occ[t] > 0 ? occ[t] : 4;
    const size_t stat = 20 * 1024;  // static LDS of k_small (histogram + scratch)
    f.refine_cap_max = (uint32_t)((budget - FINISH_STATIC_LDS) / 8);
    f.small_cap_max = (uint32_t)((budget - stat) / 16);
    // Defaults keep k_finish (many tasks) and k_small at two or more blocks per CU.
    f.refine_cap = std::min<uint32_t>(4096, f.refine_cap_max);
    f.small_cap = std::min<uint32_t>(2048, f.small_cap_max);
    // register-blocked sort: up to 64 per thread, LDS scratch of a power of two
    uint32_t sc = BLOCK;
    while (sc * 2 <= (uint32_t)(budget / 8) && sc * 2 <= (uint32_t)BLOCK * 64) sc *= 2;
    f.sort_cap = sc;
    f.device = dev;
  }
  *out = &f;
  return 0;
}

int enqueue(Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s, PhaseTimer* tm) {
  switch (o.key_type) {
    case KEY_F32: return enqueue_kf32(p, f, o, s, tm);
    case KEY_F16: return enqueue_kf16(p, f, o, s, tm);
    case KEY_BF16: return enqueue_kbf16(p, f, o, s, tm);
    case KEY_I32: return enqueue_ki32(p, f, o, s, tm);
    default: return set_error(1, "unknown key type %d", o.key_type);
  }
}

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------
Pipeline::~Pipeline() {
  if (!fixed) {
    if (h_arena) (void)hipHostFree(h_arena);
    if (d_arena) (void)hipFree(d_arena);
    if (d_hist) (void)hipFree(d_hist);
    if (d_shist) (void)hipFree(d_shist);
    if (d_cand) (void)hipFree(d_cand);
  }
  if (upload_done) (void)hipEventDestroy(upload_done);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
}

static size_t grow(size_t have, size_t need) {
  size_t c = have ? have : 16;
  while (c < need) c *= 2;
  return c;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Offsets of the arena's arrays for counts c: the uploaded part (everything
// the host builds), then the device-only control blocks.
struct ArenaLayout {
  size_t large, bases, sbases, cbases, mbases, cstart, bmap, xmap, rmap, small, med, upload, ctl, total;
};

static ArenaLayout arena_layout(const StageCounts& c) {
  ArenaLayout L;
  size_t o = 0;
  const size_t pre = (size_t)(c.nl + 1) * sizeof(int64_t);
  L.large = o; o = align256(o + (size_t)c.nl * sizeof(TaskDesc));
  L.bases = o; o = align256(o + pre);
  L.sbases = o; o = align256(o + pre);
  L.cbases = o; o = align256(o + pre);
  L.mbases = o; o = align256(o + pre);
  L.cstart = o; o = align256(o + (size_t)(c.nch + 1) * sizeof(int64_t));
  L.bmap = o; o = align256(o + (size_t)c.nbmap * sizeof(uint32_t));
  L.xmap = o; o = align256(o + (size_t)c.nsplit * sizeof(uint2));
  L.rmap = o; o = align256(o + (size_t)c.nrt * sizeof(uint2));
  L.small = o; o = align256(o + (size_t)c.ns * sizeof(TaskDesc));
  L.med = o; o = align256(o + (size_t)c.nm * sizeof(TaskDesc));
  L.upload = o;
  L.ctl = o; o = align256(o + (size_t)c.nl * sizeof(TaskCtl));
  L.total = o;
  return L;
}

int Pipeline::bind_fixed(char* dev, size_t dev_bytes, char* host, size_t host_bytes, int64_t max_long,
                         int64_t max_cand) {
  const size_t hb = align256((size_t)max_long * HIST_BINS * sizeof(uint32_t));
  const size_t cb = align256((size_t)max_cand * sizeof(uint64_t));
  if (!dev || !host || ((uintptr_t)dev & 255u) || 2 * hb + cb >= dev_bytes)
    return set_error(1, "fixed pipeline storage too small");
  fixed = true;
  d_hist = (uint32_t*)dev;
  d_shist = (uint32_t*)(dev + hb);
  d_cand = (uint64_t*)(dev + 2 * hb);
  cap_hist = (size_t)max_long;
  cap_cand = (size_t)max_cand;
  d_arena = dev + 2 * hb + cb;
  h_arena = host;
  cap_arena = std::min(dev_bytes - 2 * hb - cb, host_bytes);
  return 0;
}

int Pipeline::reserve(const StageCounts& c, int64_t tot, hipStream_t s) {
  if (!fixed) {
    if (!upload_done) RADIK_HIP(hipEventCreateWithFlags(&upload_done, hipEventDisableTiming));
    if (!side) {  // default priority, like every stream of this library
      RADIK_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
      RADIK_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      RADIK_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    // The pinned arena may still feed an in-flight copy.
    RADIK_HIP(hipEventSynchronize(upload_done));
  }
  const ArenaLayout L = arena_layout(c);
  if (L.total > cap_arena || !h_arena) {
    if (fixed)
      return set_error(2, "the launch needs %zu bytes of descriptors, the fixed storage holds %zu", L.total,
                       cap_arena);
    // Device buffers may be in use by queued work on s.
    RADIK_HIP(hipStreamSynchronize(s));
    const size_t cap = grow(cap_arena, std::max<size_t>(L.total, 4096));
    if (h_arena) RADIK_HIP(hipHostFree(h_arena));
    if (d_arena) RADIK_HIP(hipFree(d_arena));
    h_arena = nullptr;
    d_arena = nullptr;
    cap_arena = 0;
    if (hipHostMalloc((void**)&h_arena, cap, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&d_arena, cap) != hipSuccess)
      return set_error(6, "allocation of a %zu-byte descriptor arena failed", cap);
    cap_arena = cap;
  }
  h_large = (TaskDesc*)(h_arena + L.large);
  h_bases = (int64_t*)(h_arena + L.bases);
  h_sbases = (int64_t*)(h_arena + L.sbases);
  h_cbases = (int64_t*)(h_arena + L.cbases);
  h_mbases = (int64_t*)(h_arena + L.mbases);
  h_cstart = (int64_t*)(h_arena + L.cstart);
  h_bmap = (uint32_t*)(h_arena + L.bmap);
  h_xmap = (uint2*)(h_arena + L.xmap);
  h_rmap = (uint2*)(h_arena + L.rmap);
  h_small = (TaskDesc*)(h_arena + L.small);
  h_med = (TaskDesc*)(h_arena + L.med);
  d_large = (TaskDesc*)(d_arena + L.large);
  d_bases = (int64_t*)(d_arena + L.bases);
  d_sbases = (int64_t*)(d_arena + L.sbases);
  d_cbases = (int64_t*)(d_arena + L.cbases);
  d_mbases = (int64_t*)(d_arena + L.mbases);
  d_cstart = (int64_t*)(d_arena + L.cstart);
  d_bmap = (uint32_t*)(d_arena + L.bmap);
  d_xmap = (uint2*)(d_arena + L.xmap);
  d_rmap = (uint2*)(d_arena + L.rmap);
  d_small = (TaskDesc*)(d_arena + L.small);
  d_med = (TaskDesc*)(d_arena + L.med);
  d_ctl = (TaskCtl*)(d_arena + L.ctl);
  upload_bytes = L.upload;
  if ((size_t)c.nl > cap_hist) {
    if (fixed) return set_error(2, "the launch has %lld long tasks, the fixed storage holds %zu", (long long)c.nl,
                                cap_hist);
    RADIK_HIP(hipStreamSynchronize(s));
    size_t cap = grow(cap_hist, (size_t)c.nl);
    if (d_hist) RADIK_HIP(hipFree(d_hist));
    if (d_shist) RADIK_HIP(hipFree(d_shist));
    d_hist = nullptr;
    d_shist = nullptr;
    cap_hist = 0;
    if (hipMalloc((void**)&d_hist, cap * HIST_BINS * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&d_shist, cap * HIST_BINS * sizeof(uint32_t)) != hipSuccess)
      return set_error(6, "allocation of %zu task histograms failed", cap);
    RADIK_HIP(hipMemset(d_hist, 0, cap * HIST_BINS * sizeof(uint32_t)));
    RADIK_HIP(hipMemset(d_shist, 0, cap * HIST_BINS * sizeof(uint32_t)));
    cap_hist = cap;
  }
  if ((size_t)tot > cap_cand) {
    if (fixed)
      return set_error(2, "the launch needs %lld spill entries, the fixed storage holds %zu", (long long)tot,
                       cap_cand);
    RADIK_HIP(hipStreamSynchronize(s));
    // geometric growth (1.5x), so a slowly growing workload does not
    // reallocate (and synchronise) on every call
    const size_t cap = std::max<size_t>((size_t)tot, cap_cand + cap_cand / 2);
    if (d_cand) RADIK_HIP(hipFree(d_cand));
    d_cand = nullptr;
    cap_cand = 0;
    if (hipMalloc((void**)&d_cand, cap * sizeof(uint64_t)) != hipSuccess)
      return set_error(6, "allocation of a %zu-entry candidate spill buffer failed", cap);
    cap_cand = cap;
  }
  return 0;
}

// Sample groups (of SAMPLE_GROUP elements) read to predict a long task's
// first-digit bound: about 1/64 of the task, 16..1024 groups (at least a
// quarter of a task of 2^14 elements).  Tasks shorter than min_len elements
// (default 2^15; re-reading them in k_medium costs less than sampling), or
// whose k exceeds 1/64 of their length, are not predicted: k_medium,
// k_finish (or pass B) re-reads them.
static int64_t sample_groups(int64_t len, int64_t k, int64_t min_len) {
  const int64_t m = k < len ? k : len;
  if (len < min_len || len < (int64_t)SAMPLE_GROUP * 16 || m <= 0 || m * 64 > len) return 0;
  int64_t g = len / ((int64_t)SAMPLE_GROUP * 64);
  if (g < 16) g = 16;
  if (g > 1024) g = 1024;
  return g;
}

// Pass A's chunk table over `total` elements (chunks a multiple of gran, at
// least min_chunk): equal chunks of about total / blocks, or (guided) chunks
// of about the remaining elements / (spread * blocks), so the chunks taken
// last are small and the blocks finish together (a last piece under half of
// min_chunk joins its predecessor).  Writes the nch + 1 starts when `out` is
// non-null; returns nch.
static int64_t plan_chunks(int64_t total, int64_t blocks, int64_t spread, bool guided, int64_t gran,
                           int64_t min_chunk, int64_t* out) {
  int64_t n = 0, g = 0;
  const int64_t eq = (std::max((total + blocks - 1) / blocks, min_chunk) + gran - 1) / gran * gran;
  while (g < total) {
    if (out) out[n] = g;
    const int64_t rem = total - g;
    if (!guided) {
      g += std::min(eq, rem);
    } else {
      int64_t c = rem / (spread * blocks);
      if (c < min_chunk) c = min_chunk;
      c = (c + gran - 1) / gran * gran;
      g = (rem - c < min_chunk / 2) ? total : g + c;
    }
    ++n;
  }
  if (out) out[n] = total;
  return n;
}

// First task (largest t with pre[t] <= g) of each chunk start g = starts[c].
static void fill_chunk_map(uint32_t* out, int64_t nch, const int64_t* starts, const int64_t* pre, int64_t nt) {
  int64_t t = 0;
  for (int64_t c = 0; c < nch; ++c) {
    while (t + 1 < nt && pre[t + 1] <= starts[c]) ++t;
    out[c] = (uint32_t)t;
  }
}

// First task (largest t with pre[t] <= g0) of each block b, g0 = gbeg + b * chunk.
static void fill_block_map(uint32_t* out, int grid, int64_t gbeg, int64_t chunk, const int64_t* pre, int64_t nt) {
  int64_t t = 0;
  for (int b = 0; b < grid; ++b) {
    const int64_t g0 = gbeg + (int64_t)b * chunk;
    while (t + 1 < nt && pre[t + 1] <= g0) ++t;
    out[b] = (uint32_t)t;
  }
}

// Medium tasks: see MED_LEN.
static inline bool is_medium(int64_t len, int64_t k, bool pred, const LaunchOpts& o) {
  const int64_t m = k < len ? k : len;
  return o.medium && !pred && len <= MED_LEN && m <= (int64_t)MED_CAP;
}

// Short tasks run whole in one block (k_small); small_cap 0 routes every
// task through the streaming passes.
static inline bool is_short(int64_t len, uint32_t small_cap) {
  return small_cap > 0 && len <= (int64_t)small_cap;
}

// Pass A's table: with blocks_per_cu up to the resident pass-A blocks per
// CU, equal chunks for blocks_per_cu blocks per CU (one round when they all
// fit); above it, guided chunks over the resident blocks with spread
// blocks_per_cu / (2 * resident), for launches whose blocks start unevenly
// (CUs shared with the short and medium tasks' kernels).
struct ChunkShape {
  int64_t slots, spread, gran, min_chunk;
  bool guided;
};

static ChunkShape chunk_shape(const DevFacts& f, const LaunchOpts& o) {
  ChunkShape c;
  const int res_cu = std::max(1, f.hist_blocks_per_cu[o.key_type]);
  c.guided = o.blocks_per_cu > res_cu;
  c.slots = (int64_t)f.cus * (c.guided ? res_cu : std::max(1, o.blocks_per_cu));
  c.spread = c.guided ? std::max(1, o.blocks_per_cu / (2 * res_cu)) : 1;
  c.gran = (int64_t)BLOCK * (16 / key_bytes(o.key_type)) * o.unr;
  c.min_chunk = (c.guided ? 8 : 4) * c.gran;
  return c;
}

// k_split blocks of a task: one per 2^20 elements, up to 256, for tasks
// longer than BIG_LEN.
static inline int64_t split_blocks(int64_t len) { return len > BIG_LEN ? std::min<int64_t>(256, len >> 20) : 0; }

// Rank-tail runs of a task (tasks longer than BIG_LEN without WITHSCALE):
// room for a list of 4 m entries (at least 32 runs), up to RT_MAX_RUNS.  A
// task whose winners and candidates exceed it finishes in k_finish.
static inline int64_t rt_runs(int64_t len, int64_t k, uint32_t flags) {
  const int64_t m = k < len ? k : len;
  if (len <= BIG_LEN || m <= 0 || (flags & F_WITHSCALE)) return 0;
  const int64_t want = std::min<int64_t>(len, std::max<int64_t>(4 * m, 32 * (int64_t)RT_RUN));
  return std::min<int64_t>(RT_MAX_RUNS, (want + RT_RUN - 1) / RT_RUN);
}

// Long tasks are laid out predicted first, then the ones pass B always
// re-reads, so pass B can cover just the latter's elements densely.
int stage_tasks(Pipeline& p, const TaskSpec* tasks, int64_t n, const DevFacts& f, const LaunchOpts& o,
                hipStream_t s) {
  const uint32_t scap = eff_small_cap(f, o);
  StageCounts cnt;
  int64_t tot = 0, tot_med = 0, groups = 0;
  for (int64_t t = 0; t < n; ++t) {
    const int64_t len = tasks[t].len, k = tasks[t].k;
    const int64_t g = o.predict ? sample_groups(len, k, o.sample_min_len) : 0;
    if (is_short(len, scap)) {
      ++cnt.ns;
    } else if (is_medium(len, k, g > 0, o)) {
      ++cnt.nm;
      tot_med += len;
    } else {
      ++cnt.nl;
      tot += len;
      groups += g;
      cnt.nsplit += split_blocks(len);
      cnt.nrt += rt_runs(len, k, tasks[t].flags);
    }
  }
  const ChunkShape cs = chunk_shape(f, o);
  cnt.nch = cnt.nl ? plan_chunks(tot, cs.slots, cs.spread, cs.guided, cs.gran, cs.min_chunk, nullptr) : 0;
  const int64_t grid_s = groups > 0 ? (groups + SAMPLE_GROUPS_PER_BLOCK - 1) / SAMPLE_GROUPS_PER_BLOCK : 0;
  cnt.nbmap = cnt.nch + grid_s;
  int rc = p.reserve(cnt, tot + tot_med, s);
  if (rc) return rc;
  const int64_t nl = cnt.nl, ns = cnt.ns, nm = cnt.nm, nch = cnt.nch;
  int64_t il = 0, is = 0, im = 0, base = 0, mbase = tot, max_m = 0, runs = 0, msum = 0, tot_small = 0;
  int64_t max_m_med = 0;
  groups = 0;
  bool big = false, pos_bounds = false;
  // Tasks are staged longest first (o.longest_first): the one-block-per-task
  // kernels (prediction, finishing, medium tasks) then start their slowest
  // blocks first instead of whenever a slot frees up.  Results go through
  // each task's own pointers, so the order is free.
  std::vector<int64_t> ord((size_t)n);
  std::iota(ord.begin(), ord.end(), (int64_t)0);
  if (o.longest_first)
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return tasks[a].len > tasks[b].len; });
  // pass 0: predicted long tasks; pass 1: the others
  for (int pass = 0; pass < 2; ++pass)
  for (int64_t q = 0; q < n; ++q) {
    const TaskSpec& ts = tasks[ord[(size_t)q]];
    const bool shrt = is_short(ts.len, scap);
    const bool pred = o.predict && !shrt && sample_groups(ts.len, ts.k, o.sample_min_len) > 0;
    const bool med = !shrt && is_medium(ts.len, ts.k, pred, o);
    if ((pass == 0) != pred) continue;
    TaskDesc d;
    d.in = ts.in;
    d.idx_in = ts.idx_in;
    d.val_out = ts.val_out;
    d.idx_out = ts.idx_out;
    d.len = ts.len;
    d.k = (int32_t)ts.k;
    d.flags = ts.flags;
    d.nblk = 0;
    d.rt_runs = 0;
    if (shrt) {
      d.base = 0;
      tot_small += ts.len;
      p.h_small[is++] = d;
    } else if (med) {
      d.base = mbase;  // spill region behind the long tasks'
      mbase += ts.len;
      const int64_t m = ts.k < ts.len ? ts.k : ts.len;
      if (m > max_m_med) max_m_med = m;
      p.h_med[im++] = d;
    } else {
      d.base = base;
      p.h_bases[il] = base;
      p.h_sbases[il] = groups;
      const int64_t m = ts.k < ts.len ? ts.k : ts.len;
      if (m > max_m) max_m = m;
      p.h_cbases[il] = runs;
      p.h_mbases[il] = msum;
      runs += std::max<int64_t>(1, (m + SORT_CHUNK - 1) / SORT_CHUNK);  // run 0 also pads
      msum += m;
      if (ts.len > BIG_LEN) big = true;
      p.h_large[il++] = d;
      base += ts.len;
      groups += pred ? sample_groups(ts.len, ts.k, o.sample_min_len) : 0;
      if (pred && sample_groups(ts.len, ts.k, o.sample_min_len) <= PRED_MAX_G) pos_bounds = true;
    }
  }
  p.h_bases[il] = base;
  p.h_sbases[il] = groups;
  p.h_cbases[il] = runs;
  p.h_mbases[il] = msum;
  p.has_big = big;
  p.pos_bounds = pos_bounds;
  p.n_med = nm;
  p.max_m_med = (uint32_t)max_m_med;
  // Launch shapes of the streaming kernels, pass A's chunk table and the
  // chunk -> task maps; each long task learns how many chunks share it.
  if (nch > 0) plan_chunks(tot, cs.slots, cs.spread, cs.guided, cs.gran, cs.min_chunk, p.h_cstart);
  p.n_chunks = nch;
  p.grid_h = (int)nch;
  p.grid_s = (int)grid_s;
  {
    int64_t c = 0;
    for (int64_t t = 0; t < nl; ++t) {
      const int64_t b0 = p.h_bases[t], len = p.h_large[t].len;
      p.h_large[t].rt_runs = 0;
      if (len <= 0) {
        p.h_large[t].nblk = 0;
        continue;
      }
      while (c + 1 < nch && p.h_cstart[c + 1] <= b0) ++c;  // chunk holding the first element
      int64_t c2 = c;
      while (c2 + 1 < nch && p.h_cstart[c2 + 1] <= b0 + len - 1) ++c2;  // ... and the last
      p.h_large[t].nblk = (int32_t)(c2 - c + 1);
    }
  }
  p.off_h = 0;
  p.off_s = p.off_h + nch;
  p.n_bmap = cnt.nbmap;
  fill_chunk_map(p.h_bmap + p.off_h, nch, p.h_cstart, p.h_bases, nl);
  int64_t ix = 0, ir = 0, rt_max = 0;
  for (int64_t t = 0; t < nl; ++t) {
    TaskDesc& d = p.h_large[t];
    const uint32_t nb = (uint32_t)split_blocks(d.len);
    for (uint32_t j = 0; j < nb; ++j) p.h_xmap[ix++] = make_uint2((uint32_t)t, j | (nb << 16));
    const uint32_t nr = (uint32_t)rt_runs(d.len, d.k, d.flags);
    d.rt_runs = (int32_t)nr;
    for (uint32_t j = 0; j < nr; ++j) p.h_rmap[ir++] = make_uint2((uint32_t)t, j | (nr << 16));
    if ((int64_t)nr > rt_max) rt_max = nr;
  }
  p.n_split = cnt.nsplit;
  p.n_rt = cnt.nrt;
  p.rt_lds = (uint32_t)std::min<int64_t>(RT_LDS, rt_max * RT_RUN);
  fill_block_map(p.h_bmap + p.off_s, p.grid_s, 0, SAMPLE_GROUPS_PER_BLOCK, p.h_sbases, nl);
  p.total_runs = runs;
  p.total_m = msum;
  p.n_large = nl;
  p.n_small = ns;
  p.total = tot;
  p.total_med = tot_med;
  p.total_small = tot_small;
  p.total_groups = groups;
  p.max_m = (uint32_t)max_m;
  return 0;
}

// Host-logic diagnostics (no device call): the chunk table and the routing
// of tasks, as stage_tasks computes them.
int64_t diag_plan_chunks(int64_t total, int64_t blocks, int64_t spread, int guided, int64_t gran, int64_t min_chunk,
                         int64_t* out, int64_t cap) {
  const int64_t n = plan_chunks(total, blocks, spread, guided != 0, gran, min_chunk, nullptr);
  if (out && cap >= n + 1) plan_chunks(total, blocks, spread, guided != 0, gran, min_chunk, out);
  return n;
}

// cls[t]: 0 short (k_small), 1 medium (k_medium), 2 long and sampled, 3 long
// without a sample; groups[t]: sample groups.
void diag_route(const int64_t* lens, const int64_t* ks, int64_t n, uint32_t small_cap, int predict, int medium,
                int32_t* cls, int64_t* groups) {
  LaunchOpts o;
  o.predict = predict != 0;
  o.medium = medium != 0;
  for (int64_t t = 0; t < n; ++t) {
    const int64_t g = o.predict ? sample_groups(lens[t], ks[t], o.sample_min_len) : 0;
    groups[t] = 0;
    if (is_short(lens[t], small_cap)) cls[t] = 0;
    else if (is_medium(lens[t], ks[t], g > 0, o)) cls[t] = 1;
    else {
      cls[t] = g > 0 ? 2 : 3;
      groups[t] = g;
    }
  }
}

int upload(Pipeline& p, hipStream_t s) {
  // one copy of every host-built array (the arena's prefix)
  if (p.upload_bytes)
    RADIK_HIP(hipMemcpyAsync(p.d_arena, p.h_arena, p.upload_bytes, hipMemcpyHostToDevice, s));
  if (p.upload_done) RADIK_HIP(hipEventRecord(p.upload_done, s));
  return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free launches (radik_topk_noalloc): the pipeline lives in
// `scratch`, a file-scope device array (one copy per GPU, held until the
// process ends), and a static host arena.  Limits: NOALLOC_MAX_LONG long
// tasks (histograms), NOALLOC_MAX_ELEMS elements of long and medium tasks
// (spill regions), NOALLOC_MAX_TASKS tasks, NOALLOC_ARENA bytes of
// descriptors.  The device address of `scratch` is taken once per device
// with hipGetSymbolAddress (the name itself is not a device address in host
// code).  Calls are serialised by one mutex and synchronise their stream
// before returning; work another caller enqueues on the same GPU's scratch
// outside this function (none exists in this library) would race.
// ---------------------------------------------------------------------------
// Zero fill of two ranges of 32-bit words (the histograms of a pipeline that
// lives in fixed storage, cleared at the start of each of its launches).
__global__ __launch_bounds__(BLOCK) void k_zero2(uint32_t* __restrict__ a, int64_t na, uint32_t* __restrict__ b,
                                                 int64_t nb) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < na; i += stride) a[i] = 0u;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < nb; i += stride) b[i] = 0u;
}

constexpr int64_t NOALLOC_MAX_LONG = RADIK_NOALLOC_MAX_LONG;
constexpr int64_t NOALLOC_MAX_ELEMS = RADIK_NOALLOC_MAX_ELEMS;
constexpr int64_t NOALLOC_MAX_TASKS = RADIK_NOALLOC_MAX_TASKS;
constexpr size_t NOALLOC_ARENA = (size_t)2 << 20;
constexpr size_t NOALLOC_HIST = (size_t)NOALLOC_MAX_LONG * HIST_BINS * sizeof(uint32_t);
constexpr size_t NOALLOC_BYTES = 2 * NOALLOC_HIST + (size_t)NOALLOC_MAX_ELEMS * sizeof(uint64_t) + NOALLOC_ARENA;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int kNoallocDevices = 64;
std::mutex g_noalloc_mu;
char* g_noalloc_dev[kNoallocDevices];  // device address of `scratch`, per device
Pipeline g_noalloc_pipe[kNoallocDevices];
alignas(256) char g_noalloc_host[NOALLOC_ARENA];
TaskSpec g_noalloc_specs[NOALLOC_MAX_TASKS];
}  // namespace

// Runs the drop-in's pipeline for tasks described like radik_topk's
// arguments (task t at the sum of earlier lengths, outputs [k, T]) in the
// fixed storage; synchronises the stream.
int noalloc_run(float* val_out, void* idx_out, int idx_bytes, const float* val_in, const void* idx_in,
                const int64_t* task_lens, int64_t num_tasks, int64_t k, uint32_t tflags, hipStream_t s) {
  if (num_tasks > NOALLOC_MAX_TASKS)
    return set_error(RADIK_ELIMIT, "%lld tasks exceed the allocation-free limit of %lld", (long long)num_tasks,
                     (long long)NOALLOC_MAX_TASKS);
  int dev = 0;
  RADIK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocDevices) return set_error(1, "device ordinal %d out of range", dev);
  DevFacts* f = nullptr;
  int rc = dev_facts(&f);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  Pipeline& p = g_noalloc_pipe[dev];
  if (!g_noalloc_dev[dev]) {
    void* ptr = nullptr;
    RADIK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_error(5, "hipGetSymbolAddress returned no address for the scratch array");
    rc = p.bind_fixed((char*)ptr, NOALLOC_BYTES, g_noalloc_host, NOALLOC_ARENA, NOALLOC_MAX_LONG,
                      NOALLOC_MAX_ELEMS);
    if (rc) return rc;
    g_noalloc_dev[dev] = (char*)ptr;
  }
  int64_t off = 0;
  for (int64_t t = 0; t < num_tasks; ++t) {
    TaskSpec& sp = g_noalloc_specs[t];
    sp.in = val_in + off;
    sp.idx_in = (tflags & F_WITHIDXIN) ? (const void*)((const char*)idx_in + off * idx_bytes) : nullptr;
    sp.val_out = val_out + t * k;
    sp.idx_out = (void*)((char*)idx_out + t * k * idx_bytes);
    sp.len = task_lens[t];
    sp.k = k;
    sp.flags = tflags;
    off += task_lens[t];
  }
  LaunchOpts o;
  rc = stage_tasks(p, g_noalloc_specs, num_tasks, *f, o, s);
  if (rc) return rc;
  // histograms start at zero (a fill kernel: no memset node in a stream
  // that a caller may capture)
  if (p.n_large > 0) {
    const int64_t words = p.n_large * HIST_BINS;
    const int grid = (int)std::min<int64_t>((words + BLOCK - 1) / BLOCK, 2048);
    hipLaunchKernelGGL(k_zero2, dim3(grid), dim3(BLOCK), 0, s, p.d_hist, words, p.d_shist, words);
    RADIK_HIP(hipGetLastError());
  }
  rc = upload(p, s);
  if (rc) return rc;
  rc = enqueue(p, *f, o, s, nullptr);
  if (rc) return rc;
  RADIK_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace radik
