// radik_kernels.hip -- gfx950 kernels of the radix top-k path.
//
// Selection works on the 64-bit composite  C = rank_key(x) << 32 | position,
// smaller is better, so the top-k of a task are its k smallest composites and
// ties at the k-th value fall to the lowest position.  The first digit
// (rank-key bits [31:20], the reference's pass-1 digit, count_bin.jl:78-203,
// algorithm.jl:221-224) is resolved over the raw input, which is read once in
// the common case:
//
//   k_sample   first-digit histogram of an evenly spread ~1/64 sample of each
//              long task
//   k_predict  per task: a bound on composites (a first digit, narrowed to a
//              rank and position for heavy digits) that the task's m best
//              stay under with high probability
//   k_hist     pass A, the one full read: exact 4096-bin histogram per task
//              (LDS ds_add, one global atomic per non-zero bin per chunk) and,
//              in the same read, a spill of every composite under the bound
//              (per-wave LDS staging); tasks inside one chunk are selected at
//              once from the LDS histogram               ~ count_bin_ex_kernel!
//   k_select   (tasks over 2^20 elements) block scan of the 4096 bins ->
//              winner bound, candidate bin, residual k   ~ select_bin_kernel!
//   k_compact  pass B, only for tasks over 2^20 elements whose prediction was
//              missing or too tight: re-reads them, winners (digit < bin) to
//              the output slots, candidates (digit == bin) to the spill
//                                                       ~ select_candidate_ex_kernel!
//   k_split    (tasks over 2^20 elements) lists winners and candidates out of
//              the spill with several blocks per task
//   k_finish   per long task: the selection when pass A did not make it, then
//              the remaining digits (rank bits [19:0] as in
//              algorithm.jl:247-299, then position digits for exact ties) over
//              the candidates (in LDS, or bounded by a sample when there are
//              many), the winners and chosen candidates emitted and sorted in
//              registers                           ~ passes 2/3 + filter + sort
//   k_sort     per task, when the results exceed k_finish's LDS: bitonic sort
//              of the k results by (output key, position) ~ bitonic_sort!/merge_sort
//   k_small    one block per short task: everything above in LDS.
//   k_medium   one block per medium task: LDS histogram of one read, the
//              selection, then k_finish's code on a re-read.
//
// Streaming kernels (k_sample, k_hist, k_compact) split the launch's
// concatenated element space into chunks, so uneven task lengths balance at
// element granularity; a chunk that crosses task boundaries flushes per task.
// k_sample takes equal chunks; k_hist and k_compact take the chunks of a host
// table, shrinking towards its end; one chunk per block.
//
// Key types: this file is compiled once per key type (RADIK_DT: float32,
// float16, bfloat16, int32; radik_launch.hip), each copy in its own
// namespace.  Elements enter as their bit patterns; the order-preserving
// 32-bit key of a 16-bit type is its 16-bit key in the upper half (the lower
// half is zero, so the digit passes below bit 16 fix themselves without a
// pass).

#include "radik_device.h"

#ifndef RADIK_DT
#define RADIK_DT 0
#endif
#ifndef RADIK_KNS
#define RADIK_KNS kf32
#endif

namespace radik {
namespace RADIK_KNS {

constexpr int DT = RADIK_DT;
constexpr int EB = (DT == KEY_F16 || DT == KEY_BF16) ? 2 : 4;  // element bytes
constexpr int EPV = 16 / EB;                                   // elements per 16-byte load

// Order-preserving 32-bit key of an element's bits (unsigned order of keys
// == numeric order; -0 < +0 and NaNs outside +-Inf for the float types;
// RadiK.jl src/utils.jl:29-33 for float32), and its inverse.
__device__ __forceinline__ uint32_t ekey(uint32_t bits) {
  if constexpr (DT == KEY_F32) {
    return bits ^ ((uint32_t)((int32_t)bits >> 31) | 0x80000000u);
  } else if constexpr (DT == KEY_I32) {
    return bits ^ 0x80000000u;
  } else {
    const uint32_t h = bits & 0xFFFFu;
    return (h ^ ((h & 0x8000u) ? 0xFFFFu : 0x8000u)) << 16;
  }
}

__device__ __forceinline__ uint32_t ekey_inv(uint32_t k) {
  if constexpr (DT == KEY_F32) {
    return (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
  } else if constexpr (DT == KEY_I32) {
    return k ^ 0x80000000u;
  } else {
    const uint32_t h = k >> 16;
    return ((h & 0x8000u) ? (h ^ 0x8000u) : ~h) & 0xFFFFu;
  }
}

// Element i of an array of the key type, as bits.
__device__ __forceinline__ uint32_t load_elem(const void* p, int64_t i) {
  if constexpr (EB == 2) return (uint32_t)glb(reinterpret_cast<const uint16_t*>(p))[i];
  else return glb(reinterpret_cast<const uint32_t*>(p))[i];
}

__device__ __forceinline__ void store_elem(void* p, int64_t i, uint32_t bits) {
  if constexpr (EB == 2) glb(reinterpret_cast<uint16_t*>(p))[i] = (uint16_t)bits;
  else glb(reinterpret_cast<uint32_t*>(p))[i] = bits;
}

// Scaled value (WITHSCALE, float32 only): x - s; NaN keeps its own bits.
__device__ __forceinline__ float scaled(float x, float s) {
  float d = x - s;
  return (x == x) ? d : x;
}

// Rank key of an element: smaller is better.  xm = 0xFFFFFFFF for LARGEST,
// 0 otherwise; sc the task's scaler (WITHSCALE float32 tasks, else unused).
template <bool SCALE>
__device__ __forceinline__ uint32_t rank_bits(uint32_t bits, float sc, uint32_t xm) {
  if constexpr (SCALE && DT == KEY_F32) return ekey(__float_as_uint(scaled(__uint_as_float(bits), sc))) ^ xm;
  else return ekey(bits) ^ xm;
}

__device__ __forceinline__ uint32_t rank_key(uint32_t bits, float sc, uint32_t xm, bool scale) {
  return scale ? rank_bits<true>(bits, sc, xm) : rank_bits<false>(bits, sc, xm);
}

// The WITHSCALE scaler of a task: its first element, 0 if NaN/Inf
// (src/utils.jl:62-66); 0 when the task does not scale.
__device__ __forceinline__ float task_scaler(const TaskDesc& d) {
  if (DT != KEY_F32 || !(d.flags & F_WITHSCALE) || d.len <= 0) return 0.0f;
  float s = __uint_as_float(load_elem(d.in, 0));
  return (isnan(s) || isinf(s)) ? 0.0f : s;
}

__device__ __forceinline__ bool task_scales(const TaskDesc& d) {
  return DT == KEY_F32 && (d.flags & F_WITHSCALE) != 0;
}

// Largest t with bases[t] <= g (bases: T+1 prefix sums, bases[T] = total).
__device__ __forceinline__ int find_task(const int64_t* __restrict__ bases, int ntasks, int64_t g) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (bases[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// The vector body of a streaming read starts on a 128-byte boundary (after a
// scalar head of less than 128 bytes), so each wave's 1 KiB load covers
// whole cache lines: with streaming (non-temporal) loads a line shared by
// two wave loads is fetched from HBM twice (6 % more traffic on a batch of
// arbitrarily aligned tasks).
constexpr uint32_t BODY_ALIGN = 128 / EB;  // elements

__device__ __forceinline__ int64_t aligned_head_end(const void* p, int64_t a, int64_t e) {
  const int64_t mis = (int64_t)((((uintptr_t)p / EB) + (uint64_t)a) & (BODY_ALIGN - 1));
  const int64_t h = a + (((int64_t)BODY_ALIGN - mis) & (int64_t)(BODY_ALIGN - 1));
  return h < e ? h : e;
}

// Element j (0 <= j < EPV) of a 16-byte vector, as bits.
__device__ __forceinline__ uint32_t vec_elem(const u32x4& r, int j) {
  if constexpr (EB == 4) return r[j];
  else return (r[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
}

// Visits elements [s, e) of task array pp: scalar head up to 128-byte
// alignment, UNR 16-byte loads in flight per thread over the body, scalar
// tail.  f(bits, pos).
template <int UNR, typename F>
__device__ __forceinline__ void stream_range(const void* __restrict__ pp, int64_t s, int64_t e, F&& f) {
  const int64_t tid = threadIdx.x;
  int64_t a = s;
  const int64_t head_end = aligned_head_end(pp, a, e);
  for (int64_t i = a + tid; i < head_end; i += BLOCK) f(load_elem(pp, i), (uint32_t)i);
  a = head_end;
  const int64_t nvec = (e - a) / EPV;
  const RADIK_AS1 u32x4* __restrict__ pv =
      reinterpret_cast<const RADIK_AS1 u32x4*>(glb(reinterpret_cast<const char*>(pp)) + a * EB);
  int64_t v = tid;
  for (; v + (int64_t)(UNR - 1) * BLOCK < nvec; v += (int64_t)UNR * BLOCK) {
    u32x4 r[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) r[u] = __builtin_nontemporal_load(pv + v + (int64_t)u * BLOCK);
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t q = (uint32_t)(a + (v + (int64_t)u * BLOCK) * EPV);
#pragma unroll
      for (int j = 0; j < EPV; ++j) f(vec_elem(r[u], j), q + (uint32_t)j);
    }
  }
  for (; v < nvec; v += BLOCK) {
    const u32x4 r = __builtin_nontemporal_load(pv + v);
    const uint32_t q = (uint32_t)(a + v * EPV);
#pragma unroll
    for (int j = 0; j < EPV; ++j) f(vec_elem(r, j), q + (uint32_t)j);
  }
  for (int64_t i = a + nvec * EPV + tid; i < e; i += BLOCK) f(load_elem(pp, i), (uint32_t)i);
}

// Stores a result (value bits, position) into output slot `slot` of task d.
__device__ __forceinline__ void put_result(const TaskDesc& d, uint32_t slot, uint32_t bits, uint32_t pos) {
  store_elem(d.val_out, slot, bits);
  if (d.flags & F_IDX64) glb(reinterpret_cast<int64_t*>(d.idx_out))[slot] = (int64_t)pos;
  else glb(reinterpret_cast<int32_t*>(d.idx_out))[slot] = (int32_t)pos;
}

// Per-wave LDS staging of appended entries: lanes with `p` write their entry
// into the wave's stage; when the stage could overflow on the next append
// (or at the end of a segment) the wave reserves space for all of them with
// one global atomic and copies them out.  Entries are 64-bit; flush() hands
// them to a writer (index in the destination, entry).  Between appends a
// stage holds at most STAGE - 64 entries, so one append (<= 64) always fits.
constexpr uint32_t STAGE = 512;  // entries per wave

struct WaveStage {
  uint64_t* buf;  // this wave's STAGE entries in LDS
  uint32_t* n;    // this wave's fill, in LDS so lanes that sat out a divergent
                  // append still read the current value
};

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename W>
__device__ __forceinline__ void stage_flush(const WaveStage& st, uint32_t* ctr, W&& write) {
  lds_wave_sync();
  const uint32_t n = *st.n;
  if (n == 0) return;
  const uint64_t act = ballot(true);
  const uint32_t nact = (uint32_t)__builtin_popcountll(act);
  const uint32_t me = lanes_below(act);
  uint32_t base = 0;
  if (me == 0) base = atomicAdd(ctr, n);
  base = __builtin_amdgcn_readfirstlane(base);
  for (uint32_t i = me; i < n; i += nact) write(base + i, st.buf[i]);
  lds_wave_sync();
  if (me == 0) *st.n = 0;
  lds_wave_sync();
}

template <typename W>
__device__ __forceinline__ void stage_append(const WaveStage& st, bool p, uint64_t v, uint32_t* ctr, W&& write) {
  const uint64_t mk = ballot(p);
  if (!mk) return;
  lds_wave_sync();
  const uint32_t n0 = *st.n;
  if (p) st.buf[n0 + lanes_below(mk)] = v;
  const uint32_t n1 = n0 + (uint32_t)__builtin_popcountll(mk);
  lds_wave_sync();
  if (lanes_below(ballot(true)) == 0) *st.n = n1;
  lds_wave_sync();
  if (n1 > STAGE - 64) stage_flush(st, ctr, write);
}

// LDS histogram increment with a wave-level pre-merge: when every active
// lane hits the same bin (ties), one lane adds the count.
__device__ __forceinline__ void hist_add(uint32_t* lh, uint32_t bin) {
  const uint32_t b0 = __builtin_amdgcn_readfirstlane(bin);
  const uint64_t act = ballot(true);
  if (ballot(bin == b0) == act) {
    if (lanes_below(act) == 0) atomicAdd(&lh[b0], (uint32_t)__builtin_popcountll(act));
  } else {
    atomicAdd(&lh[bin], 1u);
  }
}

// Largest rank key whose first digit is <= b (b < 4096).
__device__ __forceinline__ uint32_t digit_bound(uint32_t b) {
  return (uint32_t)((((uint64_t)b + 1) << 20) - 1);
}

// Adds a block's LDS histogram into a task's global one and clears it.
__device__ __forceinline__ void flush_hist(uint32_t* __restrict__ lh, uint32_t* __restrict__ h) {
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) {
    const uint32_t c = lh[i];
    if (c) {
      atomicAdd(&h[i], c);
      lh[i] = 0;
    }
  }
}

// Block-wide AND and OR of the composites each thread saw (a, o): vb[0] and
// vb[1] in LDS, initialised by the caller to ~0 and 0 before a barrier.
__device__ __forceinline__ void block_and_or(uint64_t a, uint64_t o, unsigned long long* vb) {
  atomicAnd(&vb[0], (unsigned long long)a);
  atomicOr(&vb[1], (unsigned long long)o);
  __syncthreads();
}

// Bounds L <= U with L at most the ql-th and U at least the qh-th smallest
// (1 <= ql <= qh <= n) of n distinct composites v[0, n) in LDS that share the
// first digit `top`, found together over the narrow digits: one histogram
// per digit while both lie under one prefix, one each after they part.  The
// lower side stops at a bin whose smallest member is its quantile (L = the
// lowest composite of that bin), the upper one at a bin whose largest member
// is (U = the highest composite of that bin), so no sample entry lies in
// [L, U] beyond the ql-th .. qh-th smallest.
__device__ __forceinline__ void sample_bounds(const uint64_t* v, uint32_t n, uint32_t ql, uint32_t qh, uint32_t top,
                                              uint64_t vary, uint64_t same, uint32_t* lh8, uint32_t* wsum,
                                              uint32_t* res, uint64_t& L, uint64_t& U) {
  constexpr int shifts[7] = {44, 36, 32, 24, 16, 8, 0};
  constexpr int widths[7] = {8, 8, 4, 8, 8, 8, 8};
  uint64_t Pl = top, Ph = top;
  int sl = 52, sh = 52;
  bool dl = false, dh = false;
  for (int q = 0; q < 7 && !(dl && dh); ++q) {
    const int s = shifts[q], w = widths[q];
    const uint32_t nb = 1u << w;
    const int hs = s + w;
    if (((vary >> s) & (nb - 1)) == 0) {  // one bin
      const uint64_t fix = (same >> s) & (nb - 1);
      if (!dl) {
        Pl = (Pl << w) | fix;
        sl = s;
      }
      if (!dh) {
        Ph = (Ph << w) | fix;
        sh = s;
      }
      continue;
    }
    const bool shared = !dl && !dh && Pl == Ph;
    uint32_t c[1] = {0u};
    for (int side = 0; side < 2; ++side) {
      if (side == 0 ? dl : dh) continue;
      const uint64_t P = side == 0 ? Pl : Ph;
      if (!(shared && side == 1)) {  // (the shared histogram serves both)
        for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) lh8[i] = 0;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
          const uint64_t x = v[i];
          if ((x >> hs) == P) hist_add(lh8, (uint32_t)(x >> s) & (nb - 1));
        }
        __syncthreads();
        c[0] = (threadIdx.x < nb) ? lh8[threadIdx.x] : 0u;
      }
      block_find_bin<1>(c, side == 0 ? ql : qh, wsum, res);
      const uint32_t b = res[0], before = res[1], cnt = res[2];
      __syncthreads();  // res and lh8 are rewritten next
      if (side == 0) {
        Pl = (Pl << w) | b;
        sl = s;
        ql -= before;
        dl = ql == 1;
      } else {
        Ph = (Ph << w) | b;
        sh = s;
        qh -= before;
        dh = qh == cnt;
      }
    }
  }
  L = Pl << sl;
  U = ((Ph + 1ull) << sh) - 1ull;  // (wraps to ~0 at the top)
}

// ---------------------------------------------------------------------------
// Sampling: a predicted tasks owns G groups of SAMPLE_GROUP consecutive
// elements spread evenly over it (group j starts at j * len / G).  Blocks take
// equal runs of groups of the launch; per task the first-digit histogram of
// the sample goes to shist.
// ---------------------------------------------------------------------------

// Sample block b (of SAMPLE_GROUPS_PER_BLOCK groups; t: the task holding
// its first group, from the host): adds the groups' first digits to their
// tasks' sample histograms.
__device__ __forceinline__ void sample_block(int64_t b, int t, const TaskDesc* __restrict__ tasks,
                                             const int64_t* __restrict__ sbases, int ntasks, int64_t total_groups,
                                             uint32_t* __restrict__ shist, uint32_t* lh) {
  __shared__ int64_t gst[SAMPLE_GROUPS_PER_BLOCK];  // start of each group of the block in its task
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) lh[i] = 0;
  int64_t g0 = b * SAMPLE_GROUPS_PER_BLOCK;
  const int64_t g1 = min(g0 + (int64_t)SAMPLE_GROUPS_PER_BLOCK, total_groups);
  if (g0 >= g1) return;
  __syncthreads();
  while (true) {
    const TaskDesc d = tasks[t];
    const int64_t tb = sbases[t];
    const int64_t G = sbases[t + 1] - tb;
    const int64_t se = min(g1, tb + G);
    const float sc = task_scaler(d);
    const bool scale = task_scales(d);
    const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
    // group j of the task starts at j * len / G: one 64-bit division per
    // group, by one thread each, instead of one per load
    if (threadIdx.x < SAMPLE_GROUPS_PER_BLOCK) {
      const int64_t g = g0 + threadIdx.x;
      gst[threadIdx.x] = g < se ? (g - tb) * d.len / G : d.len;
    }
    __syncthreads();
    // all of this task's groups of the block in flight at once
    uint32_t xv[SAMPLE_GROUPS_PER_BLOCK];
#pragma unroll
    for (int u = 0; u < SAMPLE_GROUPS_PER_BLOCK; ++u) {
      const int64_t i = gst[u] + threadIdx.x;
      xv[u] = (i < d.len) ? load_elem(d.in, i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < SAMPLE_GROUPS_PER_BLOCK; ++u) {
      const int64_t i = gst[u] + threadIdx.x;
      if (i < d.len) atomicAdd(&lh[rank_key(xv[u], sc, xm, scale) >> 20], 1u);
    }
    __syncthreads();
    flush_hist(lh, shist + (size_t)t * HIST_BINS);
    if (se >= g1) break;
    g0 = se;
    ++t;
    while (t + 1 < ntasks && sbases[t + 1] <= g0) ++t;  // tasks without samples
    __syncthreads();
  }
}

__global__ __launch_bounds__(BLOCK) void k_sample(const TaskDesc* __restrict__ tasks,
                                                  const int64_t* __restrict__ sbases, int ntasks,
                                                  int64_t total_groups, uint32_t* __restrict__ shist,
                                                  const uint32_t* __restrict__ bmap) {
  __shared__ uint32_t lh[HIST_BINS];
  sample_block(blockIdx.x, (int)bmap[blockIdx.x], tasks, sbases, ntasks, total_groups, shist, lh);
}

// ---------------------------------------------------------------------------
// Prediction, one block per task; initialises the task's control block.
// From the sample histogram it picks the first digit b_pred of the q-th
// smallest sampled key, q = f*m + 4*sqrt(f*m) + 4 (f = sampled fraction,
// m = min(k, len)); then, re-reading the sampled elements of that digit
// (when the digit holds more than PRED_MIN and at most PRED_CAP of them, in
// a sample of at most PRED_MAX_G groups) and ranking them by counting, it
// takes the q-th smallest sampled composite (rank << 32 | position) as the
// bound (r_pred, p_pred), the position rounded up to the end of the gap
// after its sample group (the sample is contiguous groups, so positions are
// only known per group).  Pass A spills exactly the composites at or below
// that bound; the task holds at least m of them unless the sample is off by
// about four standard deviations, and then they include its m best (a miss
// only costs the fallback re-read).  Within a digit the bound follows
// positions too, so an exact tie across the threshold spills only about as
// many entries as needed.  Clears the sample bins.
// ---------------------------------------------------------------------------
template <int E>
__device__ __forceinline__ void reg_bitonic(uint64_t (&a)[E], uint64_t* xs);

constexpr uint32_t PRED_MIN = 32;
constexpr uint32_t PRED_CAP = 1024;
static_assert(PRED_CAP == BLOCK * 4, "the prediction ranks PRED_CAP composites, 4 per thread");

// Entry qq of the n <= BLOCK * E composites sv[0, n) in ascending order
// (register bitonic, sv its scratch) into out.  Whole block.
template <int E>
__device__ __forceinline__ void refine_pick(uint64_t* sv, uint32_t n, uint32_t qq, uint64_t& out) {
  uint64_t x[E];
#pragma unroll
  for (int l = 0; l < E; ++l) {
    const uint32_t i = threadIdx.x * (uint32_t)E + (uint32_t)l;
    x[l] = i < n ? sv[i] : ~0ull;
  }
  __syncthreads();  // sv is the sort's scratch from here
  reg_bitonic<E>(x, sv);
#pragma unroll
  for (int l = 0; l < E; ++l)
    if (threadIdx.x == qq / (uint32_t)E && (uint32_t)l == qq % (uint32_t)E) out = x[l];
}

// LDS of a prediction beside its PRED_CAP-entry array sv.
struct PredShared {
  uint32_t wsum[BLOCK / 64];
  uint32_t res[3];
  uint32_t nsel;
  unsigned long long vb[2];
  uint64_t qv;
  int64_t gst[PRED_MAX_G];  // sample group starts
};

__device__ __forceinline__ void predict_task(int t, const TaskDesc* __restrict__ tasks,
                                             const int64_t* __restrict__ sbases, uint32_t* __restrict__ shist,
                                             TaskCtl* __restrict__ ctl, PredShared& P, uint64_t* sv) {
  uint32_t* wsum = P.wsum;
  uint32_t* res = P.res;
  uint32_t& nsel = P.nsel;
  unsigned long long* vb = P.vb;
  const TaskDesc d = tasks[t];
  const int64_t G = sbases[t + 1] - sbases[t];
  const int64_t m = d.k < d.len ? (int64_t)d.k : d.len;
  if (threadIdx.x == 0) {
    res[0] = NO_BIN;
    nsel = 0;
    vb[0] = ~0ull;
    vb[1] = 0ull;
  }
  uint32_t need = 0;
  if (G > 0) {
    uint4* __restrict__ h4 = reinterpret_cast<uint4*>(shist + (size_t)t * HIST_BINS) + threadIdx.x * 4;
    uint32_t c[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 v = h4[q];
      c[4 * q + 0] = v.x;
      c[4 * q + 1] = v.y;
      c[4 * q + 2] = v.z;
      c[4 * q + 3] = v.w;
      h4[q] = make_uint4(0u, 0u, 0u, 0u);
    }
    const float S = (float)(G * SAMPLE_GROUP);
    const float mu = S / (float)d.len * (float)m;
    const float q = mu + 4.0f * sqrtf(mu) + 4.0f;
    need = (uint32_t)ceilf(q);
    if (m > 0 && (float)need <= S) block_find_bin<16>(c, need, wsum, res);
  }
  __syncthreads();
  const uint32_t b = res[0], before = res[1], cnt = res[2];
  uint32_t rp = 0, pp = 0;
  if (b < (uint32_t)HIST_BINS) {
    rp = digit_bound(b);  // the whole digit, unless refined below
    pp = 0xFFFFFFFFu;
    if (cnt > PRED_MIN && cnt <= PRED_CAP && G <= PRED_MAX_G) {
      // the sampled elements of digit b as composites (the groups of k_sample,
      // U groups in flight at a time)
      const float sc = task_scaler(d);
      const bool scale = task_scales(d);
      const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
      constexpr int U = 32;  // groups in flight per round (a hit mask bit each)
      // group starts (one 64-bit division per group; G <= PRED_MAX_G)
      if (threadIdx.x < (uint32_t)G) P.gst[threadIdx.x] = (int64_t)threadIdx.x * d.len / G;
      __syncthreads();
      for (int64_t j0 = 0; j0 < G; j0 += U) {
        uint32_t xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t i = (j0 + u < G ? P.gst[j0 + u] : d.len) + threadIdx.x;
          xv[u] = i < d.len ? load_elem(d.in, i) : 0u;
        }
        // this lane's hits as a mask; the wave reserves their slots with one
        // atomic (a scan of the lanes' counts places them)
        uint32_t hits = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t i = (j0 + u < G ? P.gst[j0 + u] : d.len) + threadIdx.x;
          const uint32_t rk = rank_key(xv[u], sc, xm, scale);
          if (i < d.len && (rk >> 20) == b) hits |= 1u << u;
        }
        const uint32_t nh = (uint32_t)__builtin_popcount(hits);
        const uint32_t incl = wave_incl_scan(nh);
        const uint32_t wtot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        uint32_t base = 0;
        if (wtot) {
          if (lane_id() == 0) base = atomicAdd(&nsel, wtot);
          base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
        }
        const uint32_t off = base + incl - nh;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if ((hits >> u) & 1u) {
            const uint32_t slot = off + (uint32_t)__builtin_popcount(hits & ((1u << u) - 1u));
            const uint32_t pos = (uint32_t)(P.gst[j0 + u] + threadIdx.x);
            if (slot < PRED_CAP) sv[slot] = ((uint64_t)rank_key(xv[u], sc, xm, scale) << 32) | pos;
          }
        }
      }
      __syncthreads();
      if (nsel == cnt && need > before && need - before <= cnt) {
        // the q-th smallest sampled composite is the qi-th of digit b: sort
        // the digit's sampled composites in registers (BLOCK * E entries for
        // the smallest E that holds them, +inf padding behind) and take
        // entry qi - 1
        const uint32_t qq = need - before - 1u;
        if (cnt <= (uint32_t)BLOCK) refine_pick<1>(sv, cnt, qq, P.qv);
        else if (cnt <= 2u * BLOCK) refine_pick<2>(sv, cnt, qq, P.qv);
        else refine_pick<4>(sv, cnt, qq, P.qv);
        __syncthreads();
        const uint64_t Ub = P.qv;
        rp = (uint32_t)(Ub >> 32);
        pp = (uint32_t)Ub;
        if (pp != 0xFFFFFFFFu) {  // up to the position before the next group's start
          int64_t j = (int64_t)pp * G / d.len;
          if (j > G - 1) j = G - 1;
          while (j > 0 && j * d.len / G > (int64_t)pp) --j;
          while (j + 1 < G && (j + 1) * d.len / G <= (int64_t)pp) ++j;
          pp = (j + 1 < G) ? (uint32_t)((j + 1) * d.len / G - 1) : 0xFFFFFFFFu;
        }
      }
    }
  }
  if (threadIdx.x == 0) {
    TaskCtl o = {};
    o.b_win = 0;
    o.b_cand = NO_BIN;
    o.b_pred = (b < (uint32_t)HIST_BINS) ? b : NO_BIN;
    o.r_pred = rp;
    o.p_pred = pp;
    ctl[t] = o;
  }
}

__global__ __launch_bounds__(BLOCK) void k_predict(const TaskDesc* __restrict__ tasks,
                                                   const int64_t* __restrict__ sbases,
                                                   uint32_t* __restrict__ shist,
                                                   TaskCtl* __restrict__ ctl) {
  __shared__ PredShared P;
  __shared__ uint64_t sv[PRED_CAP];
  predict_task(blockIdx.x, tasks, sbases, shist, ctl, P, sv);
}

// ---------------------------------------------------------------------------
// First-digit selection of one task from its complete 4096-bin histogram (c:
// this thread's 16 consecutive bins), called by the whole block.  Decides the
// winner bound, the candidate bin and the residual k, whether the spill of
// pass A holds the task's m best or the task is re-read, and whether their
// source (the spill, or winners + the candidate bin) fits k_finish's LDS
// (fin_cap entries).  b_pred / spilled: the prediction and the spill count of
// pass A.  Thread 0 writes the control block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void select_first(const TaskDesc& d, const uint32_t (&c)[16], uint32_t b_pred,
                                             uint32_t spilled, uint32_t fin_cap, TaskCtl* out,
                                             uint32_t* wsum, uint32_t* res) {
  const int64_t len = d.len;
  int64_t m = d.k < len ? (int64_t)d.k : len;
  if (threadIdx.x == 0) res[0] = NO_BIN;  // published by the barriers inside the scan
  if (m > 0 && m < len) block_find_bin<16>(c, (uint32_t)m, wsum, res);
  if (threadIdx.x == 0) {
    TaskCtl o = {};
    o.b_cand = NO_BIN;
    o.krem = 0;
    o.b_pred = b_pred;
    if (m <= 0) {
      o.b_win = 0;
    } else if (m >= len) {
      o.b_win = HIST_BINS;
    } else if (res[0] >= (uint32_t)HIST_BINS || res[1] >= (uint32_t)m || res[2] == 0 ||
               (uint32_t)m - res[1] > res[2]) {
      // The histogram does not hold the task's elements (cannot happen when
      // the buffer was zero): emit nothing rather than overrun the outputs.
      o.b_win = 0;
      m = 0;
    } else {
      const uint32_t b = res[0], before = res[1], cnt = res[2];
      const uint32_t krem = (uint32_t)m - before;
      if (krem == cnt) {
        o.b_win = b + 1;  // the whole bin is in: no candidates
      } else {
        o.b_win = b;
        o.b_cand = b;
        o.krem = krem;
        o.ncand = cnt;
      }
    }
    o.m = (uint32_t)(m > 0 ? m : 0);
    o.nwin = o.m - o.krem;
    o.win_ctr = 0;
    o.done = 1;
    // The spill is every composite at or below the predicted bound; with at
    // least m entries it holds the task's m best.  Their source for k_finish
    // is then the spill (n_sp entries), else a re-read (winners and the whole
    // candidate bin).
    if (o.m == 0) {
      o.spill_all = 1;
      o.cand_ctr = 0;
    } else if (b_pred == NO_BIN || spilled < o.m || spilled > (uint32_t)len) {
      o.spill_all = 0;
      o.cand_ctr = 0;
    } else {
      o.spill_all = 1;
      o.cand_ctr = spilled;
    }
    const uint64_t src = o.spill_all ? (uint64_t)o.cand_ctr : (uint64_t)o.nwin + o.ncand;
    o.fin = (o.m > 0 && src <= fin_cap) ? 1u : 0u;
    // The rank tail takes a task with runs from the launch whose spill holds
    // its m best, when its list (winners, then the spilled candidates) fits
    // those runs and the spill region behind the spill.
    o.rt = 0;
    if (d.rt_runs > 0 && o.spill_all && o.m > 0) {
      const uint64_t nl = (uint64_t)o.nwin + (o.ncand < o.cand_ctr ? o.ncand : o.cand_ctr);
      if (nl <= (uint64_t)d.rt_runs * RT_RUN && (uint64_t)o.cand_ctr + o.nwin + o.ncand <= (uint64_t)len) o.rt = 1;
    }
    *out = o;
  }
}

// Reads this thread's 16 bins of a task's global histogram and clears them
// (the buffer is zero again for the next launch).
__device__ __forceinline__ void take_hist(uint32_t* __restrict__ h, uint32_t (&c)[16]) {
  uint4* __restrict__ h4 = reinterpret_cast<uint4*>(h) + threadIdx.x * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = h4[q];
    c[4 * q + 0] = v.x;
    c[4 * q + 1] = v.y;
    c[4 * q + 2] = v.z;
    c[4 * q + 3] = v.w;
    h4[q] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// take_hist in two steps: this thread's 16 bins, then their clearing.
__device__ __forceinline__ void load_hist(const uint32_t* __restrict__ h, uint32_t (&c)[16]) {
  const uint4* __restrict__ h4 = reinterpret_cast<const uint4*>(h) + threadIdx.x * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = h4[q];
    c[4 * q + 0] = v.x;
    c[4 * q + 1] = v.y;
    c[4 * q + 2] = v.z;
    c[4 * q + 3] = v.w;
  }
}

__device__ __forceinline__ void clear_hist(uint32_t* __restrict__ h) {
  uint4* __restrict__ h4 = reinterpret_cast<uint4*>(h) + threadIdx.x * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) h4[q] = make_uint4(0u, 0u, 0u, 0u);
}

// The selection fields of o into a task's control block, leaving the
// counters win_ctr and list_ctr (which other blocks may be adding to).
__device__ __forceinline__ void publish_selection(TaskCtl* g, const TaskCtl& o) {
  g->b_win = o.b_win;
  g->b_cand = o.b_cand;
  g->krem = o.krem;
  g->ncand = o.ncand;
  g->nwin = o.nwin;
  g->m = o.m;
  g->cand_ctr = o.cand_ctr;
  g->spill_all = o.spill_all;
  g->fin = o.fin;
  g->done = o.done;
  g->rt = o.rt;
}

// Standalone selection, one block per task, of the tasks pass A did not
// select (those spread over several of its blocks): reads and clears the
// task's histogram.  (The pass-level diagnostics run it; a launch selects
// these tasks in k_compact or k_finish.)
__global__ __launch_bounds__(BLOCK) void k_select(const TaskDesc* __restrict__ tasks,
                                                  uint32_t* __restrict__ hist,
                                                  TaskCtl* __restrict__ ctl, uint32_t fin_cap) {
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  const int t = blockIdx.x;
  const TaskCtl prev = ctl[t];
  if (prev.done) return;  // selected at the end of pass A
  const TaskDesc d = tasks[t];
  uint32_t c[16];
  take_hist(hist + (size_t)t * HIST_BINS, c);
  select_first(d, c, prev.b_pred, prev.cand_ctr, fin_cap, &ctl[t], wsum, res);
}

// ---------------------------------------------------------------------------
// Pass A: exact first-digit histogram of every task; for predicted tasks the
// same read also spills every composite (rank << 32 | position) at or below
// the predicted bound (r_pred, p_pred) into the task's spill region.
// ---------------------------------------------------------------------------
// Pass A over elements [s, e) of task d.  Every loop below runs a trip
// count that is the same for all lanes of a wave (per-lane validity instead
// of per-lane bounds), so the fill n of the wave's stage (STAGE entries of
// LDS) is wave-uniform and lives in a scalar register; the vector body issues
// the next iteration's UNR loads before it processes the current ones, and
// tests the spill bound once per lane and iteration (the minimum rank key of
// the lane's elements) before placing any hit.
template <int UNR, bool SCALE, bool SPILL, bool POS>
__device__ __forceinline__ void pass_a_segment(const TaskDesc& d, int64_t s, int64_t e, float sc, uint32_t xm,
                                               uint32_t rp, uint32_t pp, uint32_t* lh, uint32_t* ctr,
                                               uint64_t* __restrict__ spill, uint64_t* stage) {
  static_assert((int)BODY_ALIGN <= BLOCK, "the scalar head takes one step");
  const uint32_t slen = (uint32_t)d.len;
  const uint32_t lane = lane_id();
  uint32_t n = 0;  // entries in this wave's stage
  // the stage to the task's spill region (one atomic per flush)
  auto flush = [&]() {
    lds_wave_sync();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(ctr, n);
    base = __builtin_amdgcn_readfirstlane(base);
    for (uint32_t i = lane; i < n; i += 64) {
      const uint32_t slot = base + i;
      if (slot < slen) glb(spill)[slot] = stage[i];
    }
    lds_wave_sync();
    n = 0;
  };
  // the composites of the lanes with h enter the stage (whole wave)
  auto append = [&](bool h, uint32_t rk, uint32_t pos) {
    const uint64_t mk = ballot(h);
    if (mk) {
      if (n > STAGE - 64) flush();
      if (h) stage[n + lanes_below(mk)] = ((uint64_t)rk << 32) | pos;
      n += (uint32_t)__builtin_popcountll(mk);
    }
  };
  // spilled: composites (rank << 32 | position) at or below (rp << 32 | pp);
  // without POS the bound takes every position of rank rp
  auto spills = [&](uint32_t rk, uint32_t pos) {
    if (POS) return rk < rp || (rk == rp && pos <= pp);
    (void)pos;
    return rk <= rp;
  };
  // one element per lane (ok: the lane holds one)
  auto one = [&](bool ok, int64_t i) {
    uint32_t rk = 0;
    if (ok) {
      rk = rank_bits<SCALE>(load_elem(d.in, i), sc, xm);
      atomicAdd(&lh[rk >> 20], 1u);
    }
    if (SPILL) append(ok && spills(rk, (uint32_t)i), rk, (uint32_t)i);
  };
  const int64_t tid = threadIdx.x;
  int64_t a = s;
  const int64_t head_end = aligned_head_end(d.in, a, e);
  if (a < head_end) one(a + tid < head_end, a + tid);
  a = head_end;
  const int64_t nvec = (e - a) / EPV;
  const RADIK_AS1 u32x4* __restrict__ pv =
      reinterpret_cast<const RADIK_AS1 u32x4*>(glb(reinterpret_cast<const char*>(d.in)) + a * EB);
  constexpr int NB = EPV * UNR;  // elements per lane per iteration
  int64_t vb = 0;                // first vector of the iteration (lane 0's)
  if (nvec >= (int64_t)UNR * BLOCK) {
    u32x4 r[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) r[u] = __builtin_nontemporal_load(pv + vb + tid + (int64_t)u * BLOCK);
    while (true) {
      const int64_t vn = vb + (int64_t)UNR * BLOCK;
      const bool more = vn + (int64_t)UNR * BLOCK <= nvec;
      u32x4 rn[UNR];
      if (more) {
#pragma unroll
        for (int u = 0; u < UNR; ++u) rn[u] = __builtin_nontemporal_load(pv + vn + tid + (int64_t)u * BLOCK);
      }
      uint32_t rk[NB];
      uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
#pragma unroll
        for (int j = 0; j < EPV; ++j) {
          rk[EPV * u + j] = rank_bits<SCALE>(vec_elem(r[u], j), sc, xm);
          atomicAdd(&lh[rk[EPV * u + j] >> 20], 1u);
          if (SPILL) mn = min(mn, rk[EPV * u + j]);
        }
      }
      if (SPILL && ballot(mn <= rp)) {
        const uint32_t p0 = (uint32_t)(a + (vb + tid) * EPV);  // position of this lane's first element
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const uint32_t pos = p0 + (uint32_t)((b / EPV) * BLOCK * EPV + (b % EPV));
          append(spills(rk[b], pos), rk[b], pos);
        }
      }
      vb = vn;
      if (!more) break;
      for (int u = 0; u < UNR; ++u) r[u] = rn[u];
    }
  }
  for (; vb < nvec; vb += BLOCK) {
    const int64_t v = vb + tid;
    const bool ok = v < nvec;
    u32x4 r = {0u, 0u, 0u, 0u};
    if (ok) r = __builtin_nontemporal_load(pv + v);
    const uint32_t q = (uint32_t)(a + v * EPV);
#pragma unroll
    for (int j = 0; j < EPV; ++j) {
      uint32_t rk = 0;
      if (ok) {
        rk = rank_bits<SCALE>(vec_elem(r, j), sc, xm);
        atomicAdd(&lh[rk >> 20], 1u);
      }
      if (SPILL) append(ok && spills(rk, q + (uint32_t)j), rk, q + (uint32_t)j);
    }
  }
  const int64_t t0 = a + nvec * EPV;  // fewer than EPV elements remain
  if (t0 < e) one(t0 + tid < e, t0 + tid);
  if (SPILL && n) flush();
}

// Relaxed device-scope load: reads where device-scope atomics are performed,
// so it sees every atomic update that completed before it.
__device__ __forceinline__ uint32_t load_dev(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// Pass A over one chunk [g0, g1) of the element space, task segment by task
// segment (t: the task holding g0).  SELECT: a task inside the chunk
// (nblk == 1) is selected from the block's LDS histogram right after its
// segment, without a global histogram; the others (and every task without
// SELECT) add their counts to the task's global histogram, selected after the
// pass.  A task that runs on past the chunk keeps its counts in lh (pend =
// the task, else -1) for the caller to flush, or to extend with a following
// chunk of the same task.  Ends without a barrier.
template <int UNR, bool SELECT, bool POSK>
__device__ __forceinline__ void hist_chunk(const TaskDesc* __restrict__ tasks, const int64_t* __restrict__ bases,
                                           int ntasks, int64_t g0, int64_t g1, int t, uint32_t* __restrict__ hist,
                                           TaskCtl* __restrict__ ctl, uint64_t* __restrict__ cand,
                                           uint32_t fin_cap, uint32_t* lh, uint64_t* stage, uint32_t* wsum,
                                           uint32_t* res, int& pend) {
  if (pend >= 0 && pend != t) {
    flush_hist(lh, hist + (size_t)pend * HIST_BINS);
    __syncthreads();
  }
  pend = -1;
  while (true) {
    const TaskDesc d = tasks[t];
    const int64_t tb = bases[t];
    const int64_t se = min(g1, tb + d.len);
    const float sc = task_scaler(d);
    const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
    const uint32_t bp = ctl[t].b_pred, rp = ctl[t].r_pred, pp = ctl[t].p_pred;
    const bool scale = task_scales(d);
    uint32_t* ctr = &ctl[t].cand_ctr;
    uint64_t* spill = cand + d.base;
    const int64_t s = g0 - tb, e = se - tb;
    if (bp == NO_BIN) {
      if (scale) pass_a_segment<UNR, true, false, false>(d, s, e, sc, xm, 0u, 0u, lh, ctr, spill, stage);
      else pass_a_segment<UNR, false, false, false>(d, s, e, sc, xm, 0u, 0u, lh, ctr, spill, stage);
    } else if (!POSK || pp == 0xFFFFFFFFu) {
      if (scale) pass_a_segment<UNR, true, true, false>(d, s, e, sc, xm, rp, pp, lh, ctr, spill, stage);
      else pass_a_segment<UNR, false, true, false>(d, s, e, sc, xm, rp, pp, lh, ctr, spill, stage);
    } else {
      if (scale) pass_a_segment<UNR, true, true, POSK>(d, s, e, sc, xm, rp, pp, lh, ctr, spill, stage);
      else pass_a_segment<UNR, false, true, POSK>(d, s, e, sc, xm, rp, pp, lh, ctr, spill, stage);
    }
    __syncthreads();
    if (se >= g1 && se < tb + d.len) {
      pend = t;  // continues past the chunk
    } else if (!SELECT) {
      flush_hist(lh, hist + (size_t)t * HIST_BINS);
    } else if (d.nblk <= 1) {
      // the whole task was in this chunk: select from the LDS histogram
      uint32_t c[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        c[j] = lh[threadIdx.x * 16 + j];
        lh[threadIdx.x * 16 + j] = 0;
      }
      select_first(d, c, bp, load_dev(ctr), fin_cap, &ctl[t], wsum, res);
    } else {
      flush_hist(lh, hist + (size_t)t * HIST_BINS);  // selected after the pass (k_select or k_finish)
    }
    if (se >= g1) break;
    g0 = se;
    ++t;
    while (t + 1 < ntasks && bases[t + 1] <= g0) ++t;  // skip empty tasks
    __syncthreads();
  }
}

// Pass A, one chunk of the host's chunk table per block (cstart: nch + 1
// starts; cmap: the task holding each chunk's start).  The table's chunks
// shrink towards its end and blocks start in about block order, so the
// blocks that start last (as earlier ones finish) carry the least work and
// the pass ends with the blocks finishing together.  POSK: some task's
// spill bound may hold a position (a launch without one runs the kernel
// without that code, which keeps its registers for the streaming loop).
template <int UNR, bool SELECT, bool POSK>
__global__ __launch_bounds__(BLOCK) void k_hist(const TaskDesc* __restrict__ tasks,
                                                const int64_t* __restrict__ bases, int ntasks,
                                                const int64_t* __restrict__ cstart,
                                                uint32_t* __restrict__ hist,
                                                TaskCtl* __restrict__ ctl,
                                                uint64_t* __restrict__ cand,
                                                const uint32_t* __restrict__ cmap, uint32_t fin_cap) {
  __shared__ uint32_t lh[HIST_BINS];
  __shared__ uint64_t stage_buf[BLOCK / 64][STAGE];
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) lh[i] = 0;
  const int ch = blockIdx.x;
  __syncthreads();
  int pend = -1;
  hist_chunk<UNR, SELECT, POSK>(tasks, bases, ntasks, cstart[ch], cstart[ch + 1], (int)cmap[ch], hist, ctl, cand,
                                fin_cap, lh, stage_buf[threadIdx.x >> 6], wsum, res, pend);
  if (pend >= 0) flush_hist(lh, hist + (size_t)pend * HIST_BINS);
}

// ---------------------------------------------------------------------------
// Pass B over several blocks, for tasks longer than min_len whose prediction
// was missing or too tight (k_finish re-reads shorter ones itself): winners
// (digit < bin) go to the output slots [0, nwin) as staging (value,
// position), candidates (digit == bin) to the task's spill region; with fin,
// both go to the spill region, winners first, as selection composites.
// ---------------------------------------------------------------------------
template <int UNR, bool SCALE>
__device__ __forceinline__ void pass_b_segment(const TaskDesc& d, int64_t s, int64_t e, float sc, uint32_t xm,
                                               const TaskCtl& c, uint32_t* wctr, uint32_t* cctr,
                                               uint64_t* __restrict__ cb, const WaveStage& sw,
                                               const WaveStage& scd) {
  const uint32_t bw = c.b_win, bc = c.b_cand, nwin = c.nwin;
  const uint32_t rmax = digit_bound(bc != NO_BIN ? bc : bw - 1);
  // fin: winners (as selection composites) and candidates go to the spill
  // region, winners first, for k_sort; else winners to the output slots.
  const bool fin = c.fin != 0;
  const uint32_t clen = fin ? c.ncand : (uint32_t)d.len;
  uint64_t* __restrict__ cdst = cb + (fin ? nwin : 0u);
  auto write_win = [&](uint32_t slot, uint64_t v) {
    if (slot < nwin) {
      if (fin) cb[slot] = v;
      else put_result(d, slot, (uint32_t)(v >> 32), (uint32_t)v);
    }
  };
  auto write_cand = [&](uint32_t slot, uint64_t v) {
    if (slot < clen) cdst[slot] = v;
  };
  stream_range<UNR>(d.in, s, e, [&](uint32_t x, uint32_t pos) {
    const uint32_t rk = rank_bits<SCALE>(x, sc, xm);
    const bool hit = rk <= rmax;
    if (ballot(hit)) {
      const uint32_t dg = rk >> 20;
      const uint64_t wv = ((uint64_t)(fin ? rk : x) << 32) | pos;
      stage_append(sw, hit && dg < bw, wv, wctr, write_win);
      stage_append(scd, hit && dg == bc, ((uint64_t)rk << 32) | pos, cctr, write_cand);
    }
  });
  stage_flush(sw, wctr, write_win);
  stage_flush(scd, cctr, write_cand);
}

// Rank-tail tasks whose list (winners and spilled candidates) is cut at the
// second-digit bin of the m-th best: those with candidates and a list longer
// than two runs.
__device__ __forceinline__ bool rt_cut(const TaskCtl& c) {
  return c.rt && c.b_cand != NO_BIN && (uint64_t)c.nwin + min(c.ncand, c.cand_ctr) > 2ull * RT_RUN;
}

// Pass B proper: one chunk of pass A's table per block (above).  The same
// blocks count, for the rank-tail tasks with a cut (rt_cut), the second
// digit (rank-key bits [19:8]) of their spilled candidates (first digit
// b_cand) into the task's sample-histogram slot (free after the prediction;
// the rank tail's run sort clears it): each block takes the part of the
// spill proportional to its part of the task.  k_split cuts the candidates it lists at the
// second-digit bin holding the m-th best, so the rank tail sorts about m
// entries instead of the whole candidate digit.  (The rank tail's list
// before the cut fits its runs, so a task adds at most RT_MAX_RUNS * RT_RUN
// counts, merged per wave.)
// Tasks longer than min_len (always spread over several blocks) are selected
// here, by every block of the task from its global histogram (the same
// result in each); the block holding the task's first element publishes it
// (publish_selection), read by the kernels after this one.  Pass B counts its
// winners in win_ctr and its candidates in list_ctr, which the selection
// leaves alone (zero since the prediction).
template <int UNR>
__global__ __launch_bounds__(BLOCK) void k_compact(const TaskDesc* __restrict__ tasks,
                                                   const int64_t* __restrict__ bases, int ntasks,
                                                   const int64_t* __restrict__ cstart,
                                                   TaskCtl* __restrict__ ctl,
                                                   uint64_t* __restrict__ cand,
                                                   const uint32_t* __restrict__ cmap, int64_t min_len,
                                                   uint32_t* __restrict__ shist, const uint32_t* __restrict__ hist,
                                                   uint32_t fin_cap) {
  __shared__ uint64_t stage_w[BLOCK / 64][STAGE];
  __shared__ uint64_t stage_c[BLOCK / 64][STAGE];
  __shared__ uint32_t stage_n[2][BLOCK / 64];
  __shared__ TaskCtl sel;
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  if (threadIdx.x < 2 * (BLOCK / 64)) stage_n[threadIdx.x / (BLOCK / 64)][threadIdx.x % (BLOCK / 64)] = 0;
  const WaveStage sw{stage_w[threadIdx.x >> 6], &stage_n[0][threadIdx.x >> 6]};
  const WaveStage scd{stage_c[threadIdx.x >> 6], &stage_n[1][threadIdx.x >> 6]};
  __syncthreads();
  // one chunk of pass A's table per block
  int64_t g0 = cstart[blockIdx.x];
  const int64_t g1 = cstart[blockIdx.x + 1];
  if (g0 >= g1) return;
  int t = (int)cmap[blockIdx.x];  // = find_task(bases, ntasks, g0), from the host
  while (true) {
    const TaskDesc d = tasks[t];
    const int64_t tb = bases[t];
    const int64_t se = min(g1, tb + d.len);
    TaskCtl c = ctl[t];
    if (d.len > min_len && !c.done) {
      uint32_t h[16];
      load_hist(hist + (size_t)t * HIST_BINS, h);
      select_first(d, h, c.b_pred, c.cand_ctr, fin_cap, &sel, wsum, res);
      __syncthreads();
      c = sel;
      if (g0 == tb && threadIdx.x == 0) publish_selection(&ctl[t], c);
      __syncthreads();  // (sel is rewritten by the next task's selection)
    }
    if (!c.spill_all && c.m > 0 && d.len > min_len) {
      const float sc = task_scaler(d);
      const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
      uint32_t* wctr = &ctl[t].win_ctr;
      uint32_t* cctr = &ctl[t].list_ctr;
      uint64_t* cb = cand + d.base;
      if (task_scales(d)) pass_b_segment<UNR, true>(d, g0 - tb, se - tb, sc, xm, c, wctr, cctr, cb, sw, scd);
      else pass_b_segment<UNR, false>(d, g0 - tb, se - tb, sc, xm, c, wctr, cctr, cb, sw, scd);
    } else if (rt_cut(c)) {
      const uint64_t n = c.cand_ctr;  // (<= len: the rank tail's list fits behind it)
      const uint32_t ps = (uint32_t)((uint64_t)(g0 - tb) * n / (uint64_t)d.len);
      const uint32_t pe = (uint32_t)((uint64_t)(se - tb) * n / (uint64_t)d.len);
      const uint64_t* __restrict__ spill = cand + d.base;
      uint32_t* __restrict__ h2 = shist + (size_t)t * HIST_BINS;
      for (uint32_t i = ps + threadIdx.x; i < pe; i += BLOCK) {
        const uint64_t x = spill[i];
        if ((uint32_t)(x >> 52) == c.b_cand) hist_add(h2, (uint32_t)(x >> 40) & 0xFFFu);
      }
    }
    if (se >= g1) break;
    g0 = se;
    ++t;
    while (t + 1 < ntasks && bases[t + 1] <= g0) ++t;
  }
}

// Loads of a candidate loop are batched LOADS_IN_FLIGHT per thread so one
// block keeps enough requests in flight when the candidates sit in global
// memory.
constexpr int LOADS_IN_FLIGHT = 8;

struct LdsSrc {
  const uint64_t* p;
  __device__ uint64_t operator[](uint32_t i) const { return p[i]; }
};
struct GlobalSrc {
  const uint64_t* __restrict__ p;
  __device__ uint64_t operator[](uint32_t i) const { return p[i]; }
};

// Global candidates: 16-byte loads (two entries), GLOBAL_LOADS in flight per
// thread, so one block keeps 32 KiB of requests outstanding.  A head entry
// before 16-byte alignment and an odd tail entry go to thread 0.
constexpr int GLOBAL_LOADS = 8;
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

template <typename F>
__device__ __forceinline__ void for_each_entry(GlobalSrc src, uint32_t n, F&& f) {
  const uint64_t* __restrict__ p = src.p;
  const uint32_t head = (((uintptr_t)p & 15u) != 0 && n > 0) ? 1u : 0u;
  if (head && threadIdx.x == 0) f(p[0]);
  const uint64_t* __restrict__ q = p + head;
  const uint32_t m = n - head;
  const uint32_t npair = m >> 1;
  const u64x2* __restrict__ v = reinterpret_cast<const u64x2*>(q);
  for (uint32_t b = threadIdx.x; b < npair; b += BLOCK * GLOBAL_LOADS) {
    u64x2 xs[GLOBAL_LOADS];
#pragma unroll
    for (int u = 0; u < GLOBAL_LOADS; ++u) {
      const uint32_t i = b + (uint32_t)u * BLOCK;
      xs[u] = (i < npair) ? v[i] : u64x2{0ull, 0ull};
    }
#pragma unroll
    for (int u = 0; u < GLOBAL_LOADS; ++u) {
      if (b + (uint32_t)u * BLOCK < npair) {
        f((uint64_t)xs[u].x);
        f((uint64_t)xs[u].y);
      }
    }
  }
  if ((m & 1u) && threadIdx.x == 0) f(q[m - 1]);
}

template <typename Src, typename F>
__device__ __forceinline__ void for_each_entry(Src src, uint32_t n, F&& f) {
  for (uint32_t b = threadIdx.x; b < n; b += BLOCK * LOADS_IN_FLIGHT) {
    uint64_t xs[LOADS_IN_FLIGHT];
#pragma unroll
    for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
      const uint32_t i = b + (uint32_t)u * BLOCK;
      xs[u] = (i < n) ? src[i] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < LOADS_IN_FLIGHT; ++u)
      if (b + (uint32_t)u * BLOCK < n) f(xs[u]);
  }
}

// ---------------------------------------------------------------------------
// Digit selection over composites (shared by the finishing code and k_small).
// Digits below the first: rank bits [19:8], [7:0], then position bits
// [31:20], [19:8], [7:0].  Starting from prefix P (the composite bits above
// `top` already fixed) and need = how many of the prefix-matching composites
// to take, finds the shortest prefix at which the bin holding the need-th
// smallest is taken whole.  Returns the shift s: the chosen composites are
// exactly those with (x >> s) <= P (P updated), among those matching the
// initial prefix.  `vary` has a bit set where the prefix-matching composites
// differ and `same` holds their common bits; a digit without varying bits has
// a single bin and is fixed without a pass (exact ties skip the rank digits).
// ---------------------------------------------------------------------------
template <bool NARROW = false, typename Src>
__device__ __forceinline__ int select_digits(Src src, uint32_t n, int top, uint64_t& P, uint32_t need,
                                             uint64_t vary, uint64_t same, uint32_t* lh, uint32_t* wsum,
                                             uint32_t* res) {
  // Wide digits (4096-bin histograms) for large sets; narrow ones (at most
  // 256 bins, top <= 52) for small sets already in LDS.
  constexpr int ND = NARROW ? 7 : 6;
  constexpr int shifts_w[6] = {52, 40, 32, 20, 8, 0};
  constexpr int widths_w[6] = {12, 12, 8, 12, 12, 8};
  constexpr int shifts_n[7] = {44, 36, 32, 24, 16, 8, 0};
  constexpr int widths_n[7] = {8, 8, 4, 8, 8, 8, 8};
  int s_final = top;
  for (int q = 0; q < ND; ++q) {
    const int s = NARROW ? shifts_n[q] : shifts_w[q];
    const int w = NARROW ? widths_n[q] : widths_w[q];
    if (s + w > top) continue;  // digit already fixed
    const uint32_t nb = 1u << w;
    const int hs = s + w;
    if (((vary >> s) & (nb - 1)) == 0) {  // one bin: every match carries this digit
      P = (hs >= 64 ? 0ull : (P << w)) | ((same >> s) & (nb - 1));
      s_final = s;
      continue;
    }
    for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) lh[i] = 0;
    __syncthreads();
    const uint64_t Pc = P;
    for_each_entry(src, n, [&](uint64_t x) {
      const bool match = (hs >= 64) ? true : ((x >> hs) == Pc);
      if (match) hist_add(lh, (uint32_t)(x >> s) & (nb - 1));
    });
    __syncthreads();
    if (w == 12) {
      uint32_t c[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) c[j] = lh[threadIdx.x * 16 + j];
      block_find_bin<16>(c, need, wsum, res);
    } else {
      uint32_t c[1];
      c[0] = (threadIdx.x < nb) ? lh[threadIdx.x] : 0u;
      block_find_bin<1>(c, need, wsum, res);
    }
    const uint32_t b = res[0], before = res[1], cnt = res[2];
    __syncthreads();  // res is rewritten by the next digit
    need -= before;
    P = (hs >= 64 ? 0ull : (P << w)) | b;
    s_final = s;
    if (need == cnt) break;
  }
  return s_final;
}


// ---------------------------------------------------------------------------
// Helpers of the finishing code (k_finish, k_medium): emitting winners held
// in the spill, resolving the candidates of the threshold bin digit by digit
// (in LDS when they fit) and emitting the chosen ones.
// ---------------------------------------------------------------------------

// Value of a spilled composite: the original element's bits (read back
// from the input for scaled tasks, else inverted from the rank key).
__device__ __forceinline__ uint32_t spilled_value(const TaskDesc& d, uint64_t x, uint32_t& pos) {
  pos = (uint32_t)x;
  if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
  if (task_scales(d)) return load_elem(d.in, pos);
  const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
  return ekey_inv((uint32_t)(x >> 32) ^ xm);
}

// ---------------------------------------------------------------------------
// Bitonic sort, ascending, of a[0..n) with virtual +inf padding to a power of
// two.  Every compare-exchange moves the minimum to the lower index ("flip"
// form), so padded slots never move and pairs reaching past n are skipped.
// Values must be distinct.  Callers synchronise before; ends synchronised.
// ---------------------------------------------------------------------------
template <typename Ptr>
__device__ __forceinline__ void bitonic_sort_asc(Ptr a, uint32_t n) {
  if (n < 2) return;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  const uint32_t half = p2 >> 1;
  for (uint32_t size = 2; size <= p2; size <<= 1) {
    const uint32_t hs = size >> 1;
    for (uint32_t p = threadIdx.x; p < half; p += BLOCK) {
      const uint32_t blk = p / hs, off = p % hs;
      const uint32_t lo = blk * size + off;
      const uint32_t hi = blk * size + size - 1 - off;
      if (hi < n) {
        const uint64_t x = a[lo], y = a[hi];
        if (x > y) { a[lo] = y; a[hi] = x; }
      }
    }
    __syncthreads();
    for (uint32_t st = size >> 2; st > 0; st >>= 1) {
      for (uint32_t p = threadIdx.x; p < half; p += BLOCK) {
        const uint32_t lo = 2 * p - (p & (st - 1));
        const uint32_t hi = lo + st;
        if (hi < n) {
          const uint64_t x = a[lo], y = a[hi];
          if (x > y) { a[lo] = y; a[hi] = x; }
        }
      }
      __syncthreads();
    }
  }
}

// Final order key of a result (value bits v): (ASCEND ? key : ~key) << 32 | position.
__device__ __forceinline__ uint64_t out_composite(const TaskDesc& d, uint32_t v, uint32_t pos) {
  const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
  return ((uint64_t)(ekey(v) ^ am) << 32) | pos;
}

// Writes the k output slots of task d from sorted composites (padding past m).
template <typename Src>
__device__ __forceinline__ void write_sorted(const TaskDesc& d, Src s, uint32_t m) {
  const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
  const bool withidx = (d.flags & F_WITHIDXIN) != 0;
  const bool i64 = (d.flags & F_IDX64) != 0;
  for (uint32_t i = threadIdx.x; i < (uint32_t)d.k; i += BLOCK) {
    uint32_t v = 0u;
    int64_t id = -1;
    if (i < m) {
      const uint64_t x = s[i];
      uint32_t pos = (uint32_t)x;
      if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
      v = ekey_inv((uint32_t)(x >> 32) ^ am);
      if (withidx) id = i64 ? glb(reinterpret_cast<const int64_t*>(d.idx_in))[pos]
                            : (int64_t)glb(reinterpret_cast<const int32_t*>(d.idx_in))[pos];
      else id = (int64_t)pos + 1;
    }
    store_elem(d.val_out, i, v);
    if (i64) glb(reinterpret_cast<int64_t*>(d.idx_out))[i] = id;
    else glb(reinterpret_cast<int32_t*>(d.idx_out))[i] = (int32_t)id;
  }
}

// ---------------------------------------------------------------------------
// Register-blocked bitonic sort of P2 = BLOCK * E composites, E per thread
// (thread t holds indices t*E .. t*E+E-1), ascending.  Strides below E run in
// registers, strides inside a wave through lane permutes (ds_bpermute), the
// widest (partner in another wave) through LDS scratch xs of P2 entries.  A
// cross-thread stage fetches all E partner entries before comparing any, so
// the permutes (or LDS reads) are in flight together instead of each waiting
// out the LDS latency on its own.  (DPP row shifts and the permlane swaps in
// place of the permutes measured slower here.)
// ---------------------------------------------------------------------------
template <int E>
__device__ __forceinline__ void reg_bitonic(uint64_t (&a)[E], uint64_t* xs) {
  constexpr uint32_t P2 = (uint32_t)BLOCK * E;
  const uint32_t t = threadIdx.x;
  for (uint32_t size = 2; size <= P2; size <<= 1) {
    for (uint32_t j = size >> 1; j >= (uint32_t)E; j >>= 1) {
      const uint32_t tj = j / E;  // partner thread = t ^ tj
      // j and size are at least E, so every entry of a thread lies in the
      // same half of its pair and of its size-block: one direction for all E
      const bool keep_min = (((t * E) & j) == 0) == (((t * E) & size) == 0);
      uint64_t y[E];
      if (tj < 64) {
        const int src = (int)(((t & 63u) ^ tj) << 2);  // byte address of the partner lane
#pragma unroll
        for (int l = 0; l < E; ++l) {
          const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)a[l]);
          const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src, (int)(uint32_t)(a[l] >> 32));
          y[l] = ((uint64_t)hi << 32) | lo;
        }
      } else {
#pragma unroll
        for (int l = 0; l < E; ++l) xs[t * E + l] = a[l];
        __syncthreads();
#pragma unroll
        for (int l = 0; l < E; ++l) y[l] = xs[(t ^ tj) * E + l];
        __syncthreads();
      }
#pragma unroll
      for (int l = 0; l < E; ++l) {
        const uint64_t x = a[l];
        a[l] = ((x > y[l]) == keep_min) ? y[l] : x;
      }
    }
#pragma unroll
    for (int jj = E / 2; jj > 0; jj >>= 1) {
      if ((uint32_t)jj < size) {
#pragma unroll
        for (int l = 0; l < E; ++l) {
          const int p = l ^ jj;
          if (p > l) {
            const bool asc = ((t * E + l) & size) == 0;
            const uint64_t x = a[l], y = a[p];
            const bool sw = asc ? (x > y) : (x < y);
            a[l] = sw ? y : x;
            a[p] = sw ? x : y;
          }
        }
      }
    }
  }
}

// One output slot from a sorted composite (or padding when x is the +inf pad).
__device__ __forceinline__ void write_slot(const TaskDesc& d, uint32_t i, uint64_t x, bool valid) {
  uint32_t v = 0u;
  int64_t id = -1;
  const bool i64 = (d.flags & F_IDX64) != 0;
  if (valid) {
    const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
    uint32_t pos = (uint32_t)x;
    if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
    v = ekey_inv((uint32_t)(x >> 32) ^ am);
    if (d.flags & F_WITHIDXIN)
      id = i64 ? glb(reinterpret_cast<const int64_t*>(d.idx_in))[pos]
               : (int64_t)glb(reinterpret_cast<const int32_t*>(d.idx_in))[pos];
    else
      id = (int64_t)pos + 1;
  }
  store_elem(d.val_out, i, v);
  if (i64) glb(reinterpret_cast<int64_t*>(d.idx_out))[i] = id;
  else glb(reinterpret_cast<int32_t*>(d.idx_out))[i] = (int32_t)id;
}

// Writes output slots t*E .. t*E+E-1 (thread t) from the sorted composites
// a[] (valid below m), then the slots past BLOCK*E up to k as padding.  The
// indices of a group of slots are gathered from idx_in before any of the
// group's stores: a store to the outputs may alias idx_in for the compiler,
// so a gather placed after one would wait out its own memory latency.
template <int E>
__device__ __forceinline__ void write_slots_regs(const TaskDesc& d, const uint64_t (&a)[E], uint32_t m) {
  constexpr int G = E < 16 ? E : 16;
  const uint32_t t = threadIdx.x, k = (uint32_t)d.k;
  const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
  const bool withidx = (d.flags & F_WITHIDXIN) != 0;
  const bool i64 = (d.flags & F_IDX64) != 0;
#pragma unroll
  for (int g = 0; g < E; g += G) {
    uint32_t val[G];
    int64_t id[G];
#pragma unroll
    for (int l = 0; l < G; ++l) {
      const uint32_t i = t * E + (uint32_t)(g + l);
      const uint64_t x = a[g + l];
      uint32_t pos = (uint32_t)x;
      if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
      const bool valid = i < m;
      val[l] = valid ? ekey_inv((uint32_t)(x >> 32) ^ am) : 0u;
      id[l] = -1;
      if (valid) {
        if (withidx)
          id[l] = i64 ? glb(reinterpret_cast<const int64_t*>(d.idx_in))[pos]
                      : (int64_t)glb(reinterpret_cast<const int32_t*>(d.idx_in))[pos];
        else
          id[l] = (int64_t)pos + 1;
      }
    }
#pragma unroll
    for (int l = 0; l < G; ++l) {
      const uint32_t i = t * E + (uint32_t)(g + l);
      if (i < k) {
        store_elem(d.val_out, i, val[l]);
        if (i64) glb(reinterpret_cast<int64_t*>(d.idx_out))[i] = id[l];
        else glb(reinterpret_cast<int32_t*>(d.idx_out))[i] = (int32_t)id[l];
      }
    }
  }
  for (uint32_t i = (uint32_t)BLOCK * E + t; i < k; i += BLOCK) write_slot(d, i, 0ull, false);
}

template <int E>
__device__ __forceinline__ void sort_task_regs(const TaskDesc& d, uint32_t m, uint64_t* xs) {
  const bool i64 = (d.flags & F_IDX64) != 0;
  const uint32_t t = threadIdx.x;
  uint64_t a[E];
#pragma unroll
  for (int l = 0; l < E; ++l) {
    const uint32_t i = t * E + l;
    if (i < m) {
      const uint32_t pos = i64 ? (uint32_t)glb(reinterpret_cast<const int64_t*>(d.idx_out))[i]
                               : (uint32_t)glb(reinterpret_cast<const int32_t*>(d.idx_out))[i];
      a[l] = out_composite(d, load_elem(d.val_out, i), pos);
    } else {
      a[l] = ~0ull;
    }
  }
  reg_bitonic<E>(a, xs);  // its widest stages synchronise the block: every read above is done
  write_slots_regs<E>(d, a, m);
}

// Output composite of a selection composite (rank key << 32 | position).
__device__ __forceinline__ uint64_t out_of_selection(const TaskDesc& d, uint64_t x) {
  uint32_t pos;
  const uint32_t v = spilled_value(d, x, pos);
  return out_composite(d, v, pos);
}

// Selection and output orders agree (no scaling; LARGEST with descending
// output, or smallest with ascending output): then the composites are equal.
__device__ __forceinline__ bool same_order(const TaskDesc& d) {
  const bool largest = (d.flags & F_LARGEST) != 0, ascend = (d.flags & F_ASCEND) != 0;
  return !task_scales(d) && largest != ascend;
}

// Sorts the m selection composites sel[0, m) (LDS) by output order in
// registers and writes the k output slots (padding past m).  xs: LDS scratch
// of BLOCK * E entries.
template <int E>
__device__ __forceinline__ void sort_selected_regs(const TaskDesc& d, const uint64_t* sel, uint32_t m, uint64_t* xs) {
  const uint32_t t = threadIdx.x;
  const bool same = same_order(d);
  uint64_t a[E];
#pragma unroll
  for (int l = 0; l < E; ++l) {
    const uint32_t i = t * E + l;
    a[l] = (i < m) ? (same ? sel[i] : out_of_selection(d, sel[i])) : ~0ull;
  }
  reg_bitonic<E>(a, xs);
  write_slots_regs<E>(d, a, m);
}

template <int EMAX>
__device__ __forceinline__ void sort_selected(const TaskDesc& d, const uint64_t* sel, uint32_t m, uint64_t* xs) {
  if constexpr (EMAX > 1) {
    if (m <= (uint32_t)BLOCK * (EMAX / 2)) {
      sort_selected<EMAX / 2>(d, sel, m, xs);
      return;
    }
  }
  sort_selected_regs<EMAX>(d, sel, m, xs);
}

template <int EMAX>
__device__ __forceinline__ void sort_any(const TaskDesc& d, uint32_t m, uint64_t* xs) {
  if constexpr (EMAX > 1) {
    if (m <= (uint32_t)BLOCK * (EMAX / 2)) {
      sort_any<EMAX / 2>(d, m, xs);
      return;
    }
  }
  sort_task_regs<EMAX>(d, m, xs);
}

// ---------------------------------------------------------------------------
// Final sort, one block per task, for launches whose results exceed
// k_finish's LDS: sorts the m results staged in the output slots, in
// registers + LDS up to BLOCK * E results, else by a bitonic in the task's
// (now free) spill region, which holds len >= m slots.
// ---------------------------------------------------------------------------
struct GlobalPtr {
  uint64_t* __restrict__ p;
  __device__ uint64_t& operator[](uint32_t i) const { return p[i]; }
};

// E: results per thread, chosen by the host from the launch's largest m; the
// dynamic LDS holds BLOCK * E entries.
template <int E>
__global__ __launch_bounds__(BLOCK) void k_sort(const TaskDesc* __restrict__ tasks,
                                                const TaskCtl* __restrict__ ctl,
                                                uint64_t* __restrict__ cand) {
  extern __shared__ uint64_t sb[];
  const int t = blockIdx.x;
  if (ctl[t].rt) return;  // sorted by the rank tail
  const TaskDesc d = tasks[t];
  const uint32_t m = ctl[t].m;
  const bool i64 = (d.flags & F_IDX64) != 0;
  if (m <= (uint32_t)BLOCK * E) {
    sort_any<E>(d, m, sb);
  } else {
    uint64_t* __restrict__ g = cand + d.base;
    for (uint32_t i = threadIdx.x; i < m; i += BLOCK) {
      const uint32_t pos = i64 ? (uint32_t)glb(reinterpret_cast<const int64_t*>(d.idx_out))[i]
                               : (uint32_t)glb(reinterpret_cast<const int32_t*>(d.idx_out))[i];
      g[i] = out_composite(d, load_elem(d.val_out, i), pos);
    }
    __threadfence_block();
    __syncthreads();
    bitonic_sort_asc(GlobalPtr{g}, m);
    __threadfence_block();
    __syncthreads();
    write_sorted(d, (const uint64_t*)g, m);
  }
}

// ---------------------------------------------------------------------------
// Sort of large result sets (m > 4096) over several blocks: k_sort_chunks
// sorts runs of SORT_CHUNK results per block (register bitonic) into the
// task's spill region; k_sort_rank then places every result at its final
// slot = its rank in its own run + the number of smaller results in each
// other run (binary searches, keys are distinct) and writes the outputs.
// cbases: prefix sums of the runs per task; mbases: of the results per task.
// ---------------------------------------------------------------------------

// Sort run b (xs: LDS scratch of SORT_CHUNK entries).
__device__ __forceinline__ void sort_chunk_block(uint32_t b, const TaskDesc* __restrict__ tasks,
                                                 const TaskCtl* __restrict__ ctl,
                                                 const int64_t* __restrict__ cbases, int ntasks,
                                                 uint64_t* __restrict__ cand, uint64_t* xs) {
  const int t = find_task(cbases, ntasks, b);
  if (ctl[t].rt) return;  // sorted by the rank tail
  const TaskDesc d = tasks[t];
  const uint32_t m = ctl[t].m;
  const uint32_t j0 = (uint32_t)(b - cbases[t]) * SORT_CHUNK;
  const bool i64 = (d.flags & F_IDX64) != 0;
  uint64_t a[8];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const uint32_t i = j0 + threadIdx.x * 8 + l;
    if (i < m) {
      const uint32_t pos = i64 ? (uint32_t)glb(reinterpret_cast<const int64_t*>(d.idx_out))[i]
                               : (uint32_t)glb(reinterpret_cast<const int32_t*>(d.idx_out))[i];
      a[l] = out_composite(d, load_elem(d.val_out, i), pos);
    } else {
      a[l] = ~0ull;
    }
  }
  reg_bitonic<8>(a, xs);
  uint64_t* __restrict__ g = cand + d.base;
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const uint32_t i = j0 + threadIdx.x * 8 + l;
    if (i < m) g[i] = a[l];
  }
  if (j0 == 0)  // padding slots past m
    for (uint32_t i = m + threadIdx.x; i < (uint32_t)d.k; i += BLOCK) write_slot(d, i, 0ull, false);
}

__global__ __launch_bounds__(BLOCK) void k_sort_chunks(const TaskDesc* __restrict__ tasks,
                                                       const TaskCtl* __restrict__ ctl,
                                                       const int64_t* __restrict__ cbases, int ntasks,
                                                       uint64_t* __restrict__ cand) {
  __shared__ uint64_t xs[SORT_CHUNK];
  sort_chunk_block(blockIdx.x, tasks, ctl, cbases, ntasks, cand, xs);
}

__global__ __launch_bounds__(BLOCK) void k_sort_rank(const TaskDesc* __restrict__ tasks,
                                                     const TaskCtl* __restrict__ ctl,
                                                     const int64_t* __restrict__ mbases, int ntasks,
                                                     int64_t total, const uint64_t* __restrict__ cand) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= total) return;
  const int t = find_task(mbases, ntasks, e);
  if (ctl[t].rt) return;  // sorted by the rank tail
  const TaskDesc d = tasks[t];
  const uint32_t m = ctl[t].m;
  const uint32_t j = (uint32_t)(e - mbases[t]);
  if (j >= m) return;
  const uint64_t* __restrict__ g = cand + d.base;
  const uint64_t x = g[j];
  const uint32_t own = j / SORT_CHUNK;
  const uint32_t runs = (m + SORT_CHUNK - 1) / SORT_CHUNK;
  uint32_t pos = j - own * SORT_CHUNK;
  // lower bounds of x in the other runs, eight searches in lockstep
  for (uint32_t r0 = 0; r0 < runs; r0 += 8) {
    uint32_t lo[8], hi[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t r = r0 + u;
      lo[u] = 0;
      hi[u] = (r < runs && r != own) ? min(SORT_CHUNK, m - r * SORT_CHUNK) : 0u;
    }
    for (int step = 0; step < 12; ++step) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (lo[u] < hi[u]) {
          const uint32_t mid = (lo[u] + hi[u]) >> 1;
          if (g[(r0 + u) * SORT_CHUNK + mid] < x) lo[u] = mid + 1;
          else hi[u] = mid;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) pos += lo[u];
  }
  if (pos < m) write_slot(d, pos, x, true);
}

// ---------------------------------------------------------------------------
// Finish, one block per long task, after pass A and its selection (and pass B
// over several blocks for tasks longer than big_len): gathers the task's
// winners and candidates from its spill, or re-reads the task itself when it
// has no usable spill (len <= big_len), resolves the krem best candidates
// digit by digit (rank bits [19:0], then positions for exact ties; the
// reference's passes 2/3 and filter, algorithm.jl:247-330) and sorts the m
// results into the outputs.
//   fin tasks (nwin + ncand <= cap): winners and candidates in LDS (v),
//     narrow digits; winners and chosen candidates gathered in sel and sorted
//     in registers.
//   others: winners to the output slots, candidates behind the spill (or at
//     the start of the spill region after a re-read, or left among the spill
//     when there is no room); candidates resolved in LDS when they fit cap,
//     else in global memory with wide digits (the 4096-bin histogram lives in
//     v); chosen ones emitted behind the winners; with sort_here the m output
//     slots are then sorted in registers (else k_sort follows).
// Dynamic LDS: cap entries (v), then min(max m, cap) entries (sel) when fin.
// ---------------------------------------------------------------------------
constexpr int FIN_UNR = 8;  // float4 loads in flight per thread when re-reading a task
// Minimum waves per SIMD the compiler keeps room for (register budget) in
// k_finish (E <= 8) and k_medium.
#ifndef RADIK_FIN_WAVES_PER_EU
#define RADIK_FIN_WAVES_PER_EU 8
#endif
#ifndef RADIK_MED_WAVES_PER_EU
#define RADIK_MED_WAVES_PER_EU 8
#endif
constexpr int FIN_WAVES_PER_EU = RADIK_FIN_WAVES_PER_EU;
constexpr int MED_WAVES_PER_EU = RADIK_MED_WAVES_PER_EU;

// LDS of a finishing block beside its dynamic entry array.
struct FinShared {
  uint32_t lh8[256];
  uint32_t wsum[BLOCK / 64];
  uint32_t res[3];
  uint32_t wctr, cctr, ectr;
  unsigned long long vb[2];
  TaskCtl ctl;
};

// Whether k_split left a spilled long task's winners and candidates listed
// behind its spill: fin tasks winners then candidates, others (winners went
// to the output slots) only candidates.
__device__ __forceinline__ bool split_fits(const TaskCtl& c, int64_t len) {
  const uint64_t n = c.cand_ctr, need = (c.fin || c.rt) ? (uint64_t)c.nwin + c.ncand : (uint64_t)c.ncand;
  return c.spill_all && c.m > 0 && n <= (uint64_t)len && n + need <= (uint64_t)len;
}

// Everything after the first-digit selection c of task d; see k_finish.
// split: k_split ran for the tasks longer than big_len.
template <int EMAX>
__device__ __forceinline__ void finish_task(const TaskDesc& d, const TaskCtl& c, uint64_t* __restrict__ cand,
                                            uint64_t* v, uint32_t cap, int64_t big_len, int sort_here, bool split,
                                            FinShared& S) {
  uint32_t* lh8 = S.lh8;
  uint32_t* wsum = S.wsum;
  uint32_t* res = S.res;
  uint32_t& wctr = S.wctr;
  uint32_t& cctr = S.cctr;
  uint32_t& ectr = S.ectr;
  unsigned long long* vb = S.vb;
  const uint32_t m = c.m;
  if (m == 0) {
    if (sort_here)
      for (uint32_t i = threadIdx.x; i < (uint32_t)d.k; i += BLOCK) write_slot(d, i, 0ull, false);
    return;
  }
  const uint32_t nwin = c.nwin, ncand = c.ncand, krem = c.krem, bw = c.b_win, bc = c.b_cand;
  const bool fin = c.fin != 0;
  const uint32_t len = (uint32_t)d.len;
  uint64_t* __restrict__ region = cand + d.base;
  const uint32_t n_sp = c.cand_ctr < len ? c.cand_ctr : len;  // the spill
  const bool big = d.len > big_len;
  // Where the winners and candidates are:
  //   listed  - pass B over several blocks (a long task re-read) or k_split
  //             left them listed: fin tasks winners then candidates, others
  //             only candidates (the winners are in the output slots);
  //   spilled - among the spill of pass A (winners, candidates, others);
  //   neither - the task is re-read here.
  const bool reread_done = !c.spill_all && big;
  const bool listed = reread_done || (split && big && split_fits(c, d.len));
  const uint64_t* __restrict__ lst = region + (reread_done ? 0u : n_sp);
  // (a spill need not hold the whole candidate bin: k_split lists the
  // candidates it holds, list_ctr of them; pass B counts its candidates in
  // list_ctr too)
  const uint32_t nlisted = c.list_ctr < ncand ? c.list_ctr : ncand;
  // Pass B and k_split both list fin tasks' winners first (nwin of them),
  // then the candidates.
  const uint32_t nlst = (fin ? nwin : 0u) + nlisted;
  const bool spilled = c.spill_all && !listed;
  // Candidates of non-fin tasks go to LDS when they fit, else (a re-read
  // here) to the start of the spill region.  ncl: at most that many.
  const uint32_t ncl = listed ? nlst : (spilled ? (ncand < n_sp ? ncand : n_sp) : ncand);
  const bool cand_lds = fin || ncl <= cap;
  if (threadIdx.x == 0) {
    wctr = 0;
    cctr = 0;
    ectr = 0;
    vb[0] = ~0ull;
    vb[1] = 0ull;
  }
  __syncthreads();
  uint64_t a = ~0ull, o = 0ull;
  // One composite (rank << 32 | position), by all active lanes together.
  auto take = [&](uint64_t x, bool w, bool cd) {
    const uint64_t wm = ballot(w);
    if (wm) {
      const uint32_t slot = wave_append(w, wm, &wctr);
      if (w && slot < nwin) {
        if (fin) {
          v[slot] = x;
        } else {
          uint32_t pos;
          const uint32_t val = spilled_value(d, x, pos);
          put_result(d, slot, val, pos);
        }
      }
    }
    const uint64_t cm = ballot(cd);
    if (cm) {
      if (cd) {
        a &= x;
        o |= x;
      }
      const uint32_t slot = wave_append(cd, cm, &cctr);
      if (cd && slot < ncand) {
        if (cand_lds) v[(fin ? nwin : 0u) + slot] = x;
        else glb(region)[slot] = x;  // (kept apart from the LDS store: no flat access)
      }
    }
  };
  // Many candidates (spilled or listed): bound the krem-th smallest between
  // two composites L <= U read off a sample (as pass A bounds the first
  // digit), then one pass emits the winners and the candidates below L and
  // keeps those in [L, U] in LDS; checked, else the exact digits below.
  // A re-read here whose candidates exceed LDS leaves them at the start of
  // the spill region: they are bounded the same way after the re-read.
  const bool reread_sampled = !fin && !spilled && !listed && !cand_lds && krem > 0;
  const bool sampled = (!fin && (spilled || listed) && krem > 0 && ncl > cap) || reread_sampled;
  const uint64_t* ssrc = reread_sampled ? region : (spilled ? region : lst);
  const uint32_t nss = reread_sampled ? ncand : (spilled ? n_sp : nlst);
  bool exact = true;
  auto bound_by_sample = [&]() {
    // every stride-th entry, at most `want` of them (all kept: a sample cut
    // short would cover only the source's first part, whose positions need
    // not be typical)
    const uint32_t want = cap < 4096u ? cap : 4096u;
    const uint32_t stride = (nss + want - 1) / want;
    const uint32_t ns = (nss + stride - 1) / stride;  // sampled entries j * stride, ns <= want
    for (uint32_t j0 = threadIdx.x; j0 < ns; j0 += BLOCK * LOADS_IN_FLIGHT) {
      uint64_t xs[LOADS_IN_FLIGHT];
#pragma unroll
      for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
        const uint32_t j = j0 + (uint32_t)u * BLOCK;
        xs[u] = (j < ns) ? ssrc[(size_t)j * stride] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
        const uint64_t x = xs[u];
        const bool cd = (j0 + (uint32_t)u * BLOCK < ns) && (x >> 52) == bc;
        const uint64_t cm = ballot(cd);
        if (cm) {
          const uint32_t slot = wave_append(cd, cm, &cctr);
          if (cd && slot < cap) {
            v[slot] = x;
            a &= x;
            o |= x;
          }
        }
      }
    }
    __syncthreads();
    block_and_or(a, o, vb);
    const uint32_t cs = cctr < cap ? cctr : cap;
    const float mu = (float)krem * (float)ns / (float)nss;  // (sampling rate ns / nss)
    const float dev = 4.0f * sqrtf(mu) + 4.0f;
    const int64_t qlo = (int64_t)floorf(mu - dev), qhi = (int64_t)ceilf(mu + dev);
    const uint64_t sv = (uint64_t)(vb[0] ^ vb[1]), ss = (uint64_t)vb[0];
    uint64_t L = 0ull, U = ~0ull;
    if (cs > 0) {
      const uint32_t ql = qlo < 1 ? 1u : (qlo > (int64_t)cs ? cs : (uint32_t)qlo);
      const uint32_t qh = qhi > (int64_t)cs ? cs : (qhi < 1 ? 1u : (uint32_t)qhi);
      sample_bounds(v, cs, ql, qh, bc, sv, ss, lh8, wsum, res, L, U);
      if (qlo < 1) L = 0ull;
      if (qhi > (int64_t)cs) U = ~0ull;
    }
    if (threadIdx.x == 0) {
      cctr = 0;
      vb[0] = ~0ull;
      vb[1] = 0ull;
    }
    __syncthreads();
    a = ~0ull;
    o = 0ull;
    for_each_entry(GlobalSrc{ssrc}, nss, [&](uint64_t x) {
      const uint32_t dg = (uint32_t)(x >> 52);
      const bool w = spilled && dg < bw;  // (a list holds candidates only)
      const bool lo = dg == bc && x < L;
      const bool mid = dg == bc && x >= L && x <= U;
      const uint64_t wm = ballot(w);
      if (wm) {
        const uint32_t slot = wave_append(w, wm, &wctr);
        if (w && slot < nwin) {
          uint32_t pos;
          const uint32_t val = spilled_value(d, x, pos);
          put_result(d, slot, val, pos);
        }
      }
      const uint64_t lm = ballot(lo);
      if (lm) {
        const uint32_t slot = wave_append(lo, lm, &ectr);
        if (lo && slot < krem) {
          uint32_t pos;
          const uint32_t val = spilled_value(d, x, pos);
          put_result(d, nwin + slot, val, pos);
        }
      }
      const uint64_t mm = ballot(mid);
      if (mm) {
        const uint32_t slot = wave_append(mid, mm, &cctr);
        if (mid && slot < cap) {
          v[slot] = x;
          a &= x;
          o |= x;
        }
      }
    });
    __syncthreads();
    const uint32_t below = ectr, nmid = cctr;
    if (below <= krem && below + nmid >= krem && nmid <= cap) {
      exact = false;
      block_and_or(a, o, vb);
      const uint32_t need = krem - below;
      uint64_t Pm = bc;
      int sm = 52;
      if (need > 0 && need < nmid)
        sm = select_digits<true>(LdsSrc{v}, nmid, 52, Pm, need, (uint64_t)(vb[0] ^ vb[1]), (uint64_t)vb[0], lh8,
                                 wsum, res);
      if (need > 0) {
        for (uint32_t i = threadIdx.x; i < nmid; i += BLOCK) {
          const uint64_t x = v[i];
          const bool tk = (x >> sm) <= Pm;
          const uint64_t tm = ballot(tk);
          if (tm) {
            const uint32_t slot = wave_append(tk, tm, &ectr);  // continues after the `below` ones
            if (tk && slot < krem) {
              uint32_t pos;
              const uint32_t val = spilled_value(d, x, pos);
              put_result(d, nwin + slot, val, pos);
            }
          }
        }
      }
    } else {
      // the sample misled: exact digits over the whole source, every bit varying
      a = 0ull;
      o = ~0ull;
      if (threadIdx.x == 0) ectr = 0;
    }
  };
  if (sampled && !reread_sampled) {
    // bounded below, straight from the spill or the list
  } else if (listed) {
    // fin: winners and candidates into LDS; else candidates (into LDS when
    // they fit, else they stay listed)
    for (uint32_t b0 = threadIdx.x; b0 < nlst; b0 += BLOCK * LOADS_IN_FLIGHT) {
      uint64_t xs[LOADS_IN_FLIGHT];
#pragma unroll
      for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
        const uint32_t i = b0 + (uint32_t)u * BLOCK;
        xs[u] = (i < nlst) ? lst[i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
        const uint32_t i = b0 + (uint32_t)u * BLOCK;
        if (i < nlst) {
          if (cand_lds) v[i] = xs[u];
          if (!fin || i >= nwin) {
            a &= xs[u];
            o |= xs[u];
          }
        }
      }
    }
  } else if (spilled) {
    for_each_entry(GlobalSrc{region}, n_sp, [&](uint64_t x) {
      const uint32_t dg = (uint32_t)(x >> 52);
      take(x, dg < bw, dg == bc);
    });
  } else {
    const float sc = task_scaler(d);
    const bool scale = task_scales(d);
    const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
    const uint32_t rmax = digit_bound(bc != NO_BIN ? bc : bw - 1);
    stream_range<FIN_UNR>(d.in, 0, d.len, [&](uint32_t x, uint32_t pos) {
      const uint32_t rk = rank_key(x, sc, xm, scale);
      const bool hit = rk <= rmax;
      if (ballot(hit)) {
        const uint32_t dg = rk >> 20;
        take(((uint64_t)rk << 32) | pos, hit && dg < bw, hit && dg == bc);
      }
    });
  }
  if (sampled) {
    if (reread_sampled) {
      __syncthreads();  // the candidates are in the spill region
      if (threadIdx.x == 0) {
        cctr = 0;
        vb[0] = ~0ull;
        vb[1] = 0ull;
      }
      __syncthreads();
      a = ~0ull;
      o = 0ull;
    }
    bound_by_sample();
  }
  __syncthreads();
  // candidates gathered: the list's, or those met in the spill or re-read
  const uint32_t nca = listed ? nlst - (fin ? nwin : 0u) : (cctr < ncand ? cctr : ncand);
  if (exact) block_and_or(a, o, vb);
  const uint64_t vary = (uint64_t)(vb[0] ^ vb[1]), same = (uint64_t)vb[0];
  uint64_t P = bc;
  int s = 52;
  if (fin) {
    if (krem > 0 && krem < nca)
      s = select_digits<true>(LdsSrc{v + nwin}, nca, 52, P, krem, vary, same, lh8, wsum, res);
    uint64_t* sel = v + cap;
    for (uint32_t i = threadIdx.x; i < nwin; i += BLOCK) sel[i] = v[i];
    for (uint32_t i = threadIdx.x; i < nca; i += BLOCK) {
      const uint64_t x = v[nwin + i];
      const bool tk = (x >> s) <= P;
      const uint64_t tm = ballot(tk);
      if (tm) {
        const uint32_t slot = wave_append(tk, tm, &ectr);
        if (tk && slot < krem) sel[nwin + slot] = x;
      }
    }
    __syncthreads();  // sel complete; v is free for the sort's scratch
    sort_selected<EMAX>(d, sel, m, v);
    return;
  }
  if (krem > 0 && exact) {
    // the candidates: in LDS, listed, in the spill region after a re-read
    // here, or (after a misleading sample) in its source
    const uint64_t* __restrict__ src = sampled ? ssrc : (listed ? lst : region);
    const uint32_t nsrc = sampled ? nss : nca;
    const bool in_lds = cand_lds && !sampled;
    if (krem < nsrc) {
      if (in_lds) s = select_digits<true>(LdsSrc{v}, nsrc, 52, P, krem, vary, same, lh8, wsum, res);
      else s = select_digits<false>(GlobalSrc{src}, nsrc, 52, P, krem, vary, same, (uint32_t*)v, wsum, res);
    }
    auto emit = [&](uint64_t x) {
      const bool tk = (x >> 52) == bc && (x >> s) <= P;
      const uint64_t tm = ballot(tk);
      if (tm) {
        const uint32_t slot = wave_append(tk, tm, &ectr);
        if (tk && slot < krem) {
          uint32_t pos;
          const uint32_t val = spilled_value(d, x, pos);
          put_result(d, nwin + slot, val, pos);
        }
      }
    };
    if (in_lds) for_each_entry(LdsSrc{v}, nsrc, emit);
    else for_each_entry(GlobalSrc{src}, nsrc, emit);
  }
  if (sort_here) {
    __threadfence_block();
    __syncthreads();  // the block's output-slot writes, and v free
    sort_any<EMAX>(d, m, v);
  }
}

__device__ __forceinline__ void rt_runs_block(uint32_t b, const TaskDesc* __restrict__ tasks,
                                              const TaskCtl* __restrict__ ctl, uint64_t* __restrict__ cand,
                                              const uint2* __restrict__ rmap, uint32_t* __restrict__ shist,
                                              uint64_t* xs);

// Blocks [0, nfin): one per long task; blocks from nfin on: the rank tail's
// run sorts (rmap, shist), which depend on the same kernels as the finishing.
template <int EMAX>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(EMAX <= 8 ? FIN_WAVES_PER_EU : 1))) void k_finish(const TaskDesc* __restrict__ tasks,
                                                  TaskCtl* __restrict__ ctl,
                                                  uint32_t* __restrict__ hist,
                                                  uint64_t* __restrict__ cand, uint32_t cap, uint32_t fin_cap,
                                                  int64_t big_len, int sort_here, int split, int nfin,
                                                  const uint2* __restrict__ rmap, uint32_t* __restrict__ shist) {
  extern __shared__ uint64_t v[];
  __shared__ FinShared S;
  if ((int)blockIdx.x >= nfin) {  // (v holds at least RT_RUN entries)
    rt_runs_block(blockIdx.x - (uint32_t)nfin, tasks, ctl, cand, rmap, shist, v);
    return;
  }
  const int t = blockIdx.x;
  const TaskDesc d = tasks[t];
  TaskCtl c = ctl[t];
  if (!c.done) {  // spread over several pass-A blocks: select from the global histogram
    // (read only here: the histograms of the tasks pass A selected are all
    // zero, and reading them in every block costs the launch a burst of
    // 16 KiB per task)
    uint32_t h[16];
    take_hist(hist + (size_t)t * HIST_BINS, h);
    select_first(d, h, c.b_pred, c.cand_ctr, fin_cap, &S.ctl, S.wsum, S.res);
    __syncthreads();
    c = S.ctl;
    if (threadIdx.x == 0) ctl[t] = c;  // k_sort reads m
  } else if (split && d.len > big_len) {
    clear_hist(hist + (size_t)t * HIST_BINS);  // (selected in k_compact, which only reads it)
  }
  if (c.rt) return;  // finished by the rank tail
  finish_task<EMAX>(d, c, cand, v, cap, big_len, sort_here, split != 0, S);
}

// ---------------------------------------------------------------------------
// Part j of nb of a task's spill of n entries for a block of k_split: at
// least one full tile of BLOCK * LOADS_IN_FLIGHT entries per block, so a
// short spill takes few blocks.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void spill_part(uint32_t n, uint32_t j, uint32_t nb, uint32_t& s, uint32_t& e) {
  const uint32_t per = max((n + nb - 1) / nb, (uint32_t)BLOCK * LOADS_IN_FLIGHT);
  s = (uint32_t)min((uint64_t)n, (uint64_t)j * per);
  e = (uint32_t)min((uint64_t)n, (uint64_t)s + per);
}

// ---------------------------------------------------------------------------
// Split of the spills of long tasks (longer than big_len) over several blocks
// ahead of k_finish, so one block does not read a long spill alone: winners
// and candidates listed behind the spill (split_fits; fin tasks winners then
// candidates, others candidates only with the winners in the output slots),
// placed per tile of BLOCK * LOADS_IN_FLIGHT entries by one block scan and
// one global atomic per list.  Rank-tail tasks list only the candidates up
// to the second-digit bin (counted by k_compact) holding the m-th best.
// xmap: task, then block index in the task (low 16 bits) and blocks of the
// task (high 16 bits), of each block.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_split(const TaskDesc* __restrict__ tasks, TaskCtl* __restrict__ ctl,
                                                 uint64_t* __restrict__ cand, const uint2* __restrict__ xmap,
                                                 const uint32_t* __restrict__ shist) {
  constexpr int U = LOADS_IN_FLIGHT;
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t slot0[2];
  __shared__ uint32_t res[3];
  const uint2 tj = xmap[blockIdx.x];
  const uint32_t t = tj.x, j = tj.y & 0xFFFFu, nb = tj.y >> 16;
  const TaskCtl c = ctl[t];
  const TaskDesc d = tasks[t];
  if (!split_fits(c, d.len)) return;
  const uint32_t n = c.cand_ctr;
  uint32_t s, e;
  spill_part(n, j, nb, s, e);
  if (s >= e) return;
  const bool fin = c.fin != 0 || c.rt != 0;  // winners listed too (as composites)
  const uint32_t bw = c.b_win, bc = c.b_cand, nwin = c.nwin, ncand = c.ncand;
  uint32_t rl = 0xFFFFFFFFu;  // candidates listed: rank keys up to rl
  if (rt_cut(c)) {
    // the second-digit bin of the krem-th spilled candidate (the spill is a
    // prefix of the task's composite order holding its m best, so its
    // candidates up to that bin are too, with at least krem of them)
    uint32_t c2[16];
    load_hist(shist + (size_t)t * HIST_BINS, c2);
    if (threadIdx.x == 0) res[0] = NO_BIN;
    block_find_bin<16>(c2, c.krem, wsum, res);
    if (res[0] < (uint32_t)HIST_BINS) rl = (bc << 20) | (res[0] << 8) | 0xFFu;
  }
  uint64_t* __restrict__ spill = cand + d.base;
  uint64_t* __restrict__ wdst = spill + n;
  uint64_t* __restrict__ cdst = spill + n + (fin ? nwin : 0u);
  for (uint32_t tile = s; tile < e; tile += (uint32_t)BLOCK * U) {
    uint64_t xs[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = tile + (uint32_t)u * BLOCK + threadIdx.x;
      xs[u] = (i < e) ? spill[i] : 0ull;
    }
    uint32_t wmask = 0, cmask = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (tile + (uint32_t)u * BLOCK + threadIdx.x >= e) continue;
      const uint32_t dg = (uint32_t)(xs[u] >> 52);
      if (dg < bw) wmask |= 1u << u;
      else if (dg == bc && (uint32_t)(xs[u] >> 32) <= rl) cmask |= 1u << u;
    }
    const uint32_t mine = (uint32_t)__builtin_popcount(wmask) | ((uint32_t)__builtin_popcount(cmask) << 16);
    const uint32_t incl = block_incl_scan(mine, wsum);
    if (threadIdx.x == BLOCK - 1) {
      const uint32_t tw = incl & 0xFFFFu, tc = incl >> 16;
      slot0[0] = tw ? atomicAdd(&ctl[t].win_ctr, tw) : 0u;
      slot0[1] = tc ? atomicAdd(&ctl[t].list_ctr, tc) : 0u;
    }
    __syncthreads();
    const uint32_t excl = incl - mine;
    uint32_t ws = slot0[0] + (excl & 0xFFFFu), cs = slot0[1] + (excl >> 16);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (wmask & (1u << u)) {
        if (ws < nwin) {
          if (fin) {
            wdst[ws] = xs[u];
          } else {
            uint32_t pos;
            const uint32_t v = spilled_value(d, xs[u], pos);
            put_result(d, ws, v, pos);
          }
        }
        ++ws;
      } else if (cmask & (1u << u)) {
        if (cs < ncand) cdst[cs] = xs[u];
        ++cs;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rank tail, for tasks over 2^20 elements whose spill holds their m best
// (TaskCtl::rt, set at the selection): k_split has listed the task's winners
// and spilled candidates as composites, list L = [nwin winners | candidates]
// behind the spill.  The m best of the task are the m smallest of L.
//   runs       one block per run of RT_RUN entries of L (extra blocks of
//              k_finish): sorts it in place (register bitonic)
//   k_rt_rank  one block per RT_ENTRIES entries of L: the entry's rank r
//              over L (its place in its own run plus a binary search in
//              every other run, the runs staged through LDS, the block's
//              waves taking every fourth run each) selects it (r < m) and gives
//              its output slot: r itself when the output order is the
//              selection order, else the slot in the reversed order of rank
//              keys with equal keys by position (their group bounds come from
//              two more searches in the runs where x's rank key recurs).
// Both take the whole chip for one task, where k_finish would sort the m
// results in one block.
// ---------------------------------------------------------------------------
// Loads in flight per thread while a rank block fills its LDS with runs, and
// runs searched together.
constexpr int RT_FILL = 16;
constexpr int RT_ILP = 4;
// List entries per rank block (one per lane; its waves split the runs).
constexpr uint32_t RT_ENTRIES = 64;

__device__ __forceinline__ uint32_t rt_list_len(const TaskCtl& c) {
  return c.nwin + (c.list_ctr < c.ncand ? c.list_ctr : c.ncand);
}

// Run b of the rank tail (rmap[b]: task, run index): sorts the run of the
// task's list in place (xs: LDS scratch of RT_RUN entries).  Run by extra
// blocks of k_finish, beside the finishing of the other tasks.
__device__ __forceinline__ void rt_runs_block(uint32_t b, const TaskDesc* __restrict__ tasks,
                                              const TaskCtl* __restrict__ ctl, uint64_t* __restrict__ cand,
                                              const uint2* __restrict__ rmap, uint32_t* __restrict__ shist,
                                              uint64_t* xs) {
  constexpr int E = (int)(RT_RUN / BLOCK);
  const uint2 tj = rmap[b];
  const uint32_t t = tj.x, j = tj.y & 0xFFFFu;
  const TaskCtl c = ctl[t];
  if (!c.rt) return;
  if (j == 0 && rt_cut(c)) clear_hist(shist + (size_t)t * HIST_BINS);  // k_compact's, for the next launch
  const uint32_t nl = rt_list_len(c);
  const uint32_t s0 = j * RT_RUN;
  if (s0 >= nl) return;
  uint64_t* __restrict__ L = cand + tasks[t].base + c.cand_ctr;
  uint64_t a[E];
#pragma unroll
  for (int l = 0; l < E; ++l) {
    const uint32_t i = s0 + threadIdx.x * E + (uint32_t)l;
    a[l] = i < nl ? L[i] : ~0ull;
  }
  reg_bitonic<E>(a, xs);
#pragma unroll
  for (int l = 0; l < E; ++l) {
    const uint32_t i = s0 + threadIdx.x * E + (uint32_t)l;
    if (i < nl) L[i] = a[l];
  }
}

// Entries of the sorted run `run` (rl of them) below v.
__device__ __forceinline__ uint32_t run_below(const uint64_t* run, uint32_t rl, uint64_t v) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t step = RT_RUN; step > 0; step >>= 1)
    if (pos + step <= rl && run[pos + step - 1] < v) pos += step;
  return pos;
}

// Partial counts of a rank block's waves 1..3 (rank, glo, ghi per entry).
struct RankParts {
  uint32_t v[3][BLOCK / 64 - 1][RT_ENTRIES];
};

// Rank block b (RT_RUN / RT_ENTRIES of them per run of rmap): lr holds
// lds_entries of the list at a time (a multiple of RT_RUN).
__device__ __forceinline__ void rt_rank_block(uint32_t b, const TaskDesc* __restrict__ tasks,
                                              const TaskCtl* __restrict__ ctl, const uint64_t* __restrict__ cand,
                                              const uint2* __restrict__ rmap, uint32_t lds_entries, uint64_t* lr,
                                              RankParts& P) {
  constexpr uint32_t PARTS = RT_RUN / RT_ENTRIES;
  constexpr uint32_t WAVES = BLOCK / 64;
  const uint2 tj = rmap[b / PARTS];
  const uint32_t t = tj.x, j = tj.y & 0xFFFFu;
  const TaskCtl c = ctl[t];
  if (!c.rt) return;
  const TaskDesc d = tasks[t];
  const uint32_t nl = rt_list_len(c), m = c.m;
  const uint32_t q0 = j * RT_RUN + (b % PARTS) * RT_ENTRIES;
  if (q0 == 0)  // the task's first block pads the slots past m
    for (uint32_t i = m + threadIdx.x; i < (uint32_t)d.k; i += BLOCK) write_slot(d, i, 0ull, false);
  if (q0 >= nl) return;
  const uint64_t* __restrict__ L = cand + d.base + c.cand_ctr;
  // every wave holds the block's RT_ENTRIES entries, one per lane, and
  // searches every WAVES-th run
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t qi = q0 + (threadIdx.x & 63u);
  const bool valid = qi < nl;
  const uint64_t x = valid ? L[qi] : ~0ull;
  const uint32_t own = qi / RT_RUN, runs = (nl + RT_RUN - 1) / RT_RUN;
  const bool same = same_order(d);
  const uint32_t rk = (uint32_t)(x >> 32);
  const uint64_t key_lo = (uint64_t)rk << 32, key_hi = key_lo | 0xFFFFFFFFull;  // rank key rk: [key_lo, key_hi)
  uint32_t rank = 0, glo = 0, ghi = 0;  // entries below x; with a rank key below rk; at most rk
  const uint32_t per = lds_entries / RT_RUN;
  for (uint32_t r0 = 0; r0 < runs; r0 += per) {
    const uint32_t r1 = min(runs, r0 + per);
    const uint32_t e0 = r0 * RT_RUN, e1 = min(nl, r1 * RT_RUN);
    __syncthreads();  // the previous group's searches are done
    // RT_FILL loads in flight per thread before their LDS stores
    for (uint32_t i0 = e0; i0 < e1; i0 += (uint32_t)BLOCK * RT_FILL) {
      uint64_t tmp[RT_FILL];
#pragma unroll
      for (int u = 0; u < RT_FILL; ++u) {
        const uint32_t i = i0 + (uint32_t)u * BLOCK + threadIdx.x;
        tmp[u] = i < e1 ? L[i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < RT_FILL; ++u) {
        const uint32_t i = i0 + (uint32_t)u * BLOCK + threadIdx.x;
        if (i < e1) lr[i - e0] = tmp[u];
      }
    }
    __syncthreads();
    for (uint32_t r = r0 + w; r < r1; r += WAVES * RT_ILP) {
      // the binary searches of RT_ILP runs interleaved, step by step, so
      // their LDS reads are in flight together
      // (branch-free: a run past r1 has length 0 and probes entry 0)
      uint32_t lbs[RT_ILP], rls[RT_ILP], bas[RT_ILP];
#pragma unroll
      for (int q = 0; q < RT_ILP; ++q) {
        const uint32_t rr = r + (uint32_t)q * WAVES;
        const bool ok = rr < r1;
        rls[q] = ok ? min(RT_RUN, nl - rr * RT_RUN) : 0u;
        bas[q] = ok ? (rr - r0) * RT_RUN : 0u;
        lbs[q] = 0;
      }
#pragma unroll
      for (uint32_t step = RT_RUN; step > 0; step >>= 1) {
#pragma unroll
        for (int q = 0; q < RT_ILP; ++q) {
          const uint32_t pr = lbs[q] + step;
          const bool in = pr <= rls[q];
          const uint64_t y = lr[bas[q] + (in ? pr - 1u : 0u)];
          lbs[q] = (in && y < x) ? pr : lbs[q];
        }
      }
#pragma unroll
      for (int q = 0; q < RT_ILP; ++q) {
        const uint32_t rr = r + (uint32_t)q * WAVES;
        if (rr >= r1) break;
        const uint64_t* run = lr + (rr - r0) * RT_RUN;
        const uint32_t rl = min(RT_RUN, nl - rr * RT_RUN);
        const bool mine = valid && rr == own;
        const uint32_t lb = mine ? qi - rr * RT_RUN : lbs[q];
        rank += lb;
        if (!same) {
          // entries of this run with rank key rk lie next to x's place
          const uint32_t nx = mine ? lb + 1 : lb;
          const bool tie = (lb > 0 && (uint32_t)(run[lb - 1] >> 32) == rk) ||
                           (nx < rl && (uint32_t)(run[nx] >> 32) == rk);
          if (tie) {
            glo += run_below(run, rl, key_lo);
            ghi += run_below(run, rl, key_hi);  // (x < key_hi: positions stay below 2^32 - 1)
          } else {
            glo += lb;
            ghi += nx;
          }
        }
      }
    }
  }
  // the waves' partial counts to wave 0
  if (w > 0) {
    P.v[0][w - 1][threadIdx.x & 63u] = rank;
    P.v[1][w - 1][threadIdx.x & 63u] = glo;
    P.v[2][w - 1][threadIdx.x & 63u] = ghi;
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (uint32_t v = 0; v + 1 < WAVES; ++v) {
      rank += P.v[0][v][threadIdx.x];
      glo += P.v[1][v][threadIdx.x];
      ghi += P.v[2][v][threadIdx.x];
    }
    if (valid && rank < m) {
      const uint32_t slot = same ? rank : m - min(m, ghi) + (rank - glo);
      write_slot(d, slot, same ? x : out_of_selection(d, x), true);
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_rt_rank(const TaskDesc* __restrict__ tasks,
                                                   const TaskCtl* __restrict__ ctl, const uint64_t* __restrict__ cand,
                                                   const uint2* __restrict__ rmap, uint32_t lds_entries) {
  extern __shared__ uint64_t lr[];  // lds_entries, a multiple of RT_RUN
  __shared__ RankParts P;
  rt_rank_block(blockIdx.x, tasks, ctl, cand, rmap, lds_entries, lr, P);
}

// k_rt_rank's blocks [0, nrank) and k_sort_chunks' after them in one launch
// (both follow k_finish): lr holds max(lds_entries, SORT_CHUNK) entries.
__global__ __launch_bounds__(BLOCK) void k_rank_sort(const TaskDesc* __restrict__ tasks,
                                                     const TaskCtl* __restrict__ ctl, uint64_t* __restrict__ cand,
                                                     const uint2* __restrict__ rmap, uint32_t lds_entries,
                                                     uint32_t nrank, const int64_t* __restrict__ cbases,
                                                     int ntasks) {
  extern __shared__ uint64_t lr[];
  __shared__ RankParts P;
  if (blockIdx.x < nrank) rt_rank_block(blockIdx.x, tasks, ctl, cand, rmap, lds_entries, lr, P);
  else sort_chunk_block(blockIdx.x - nrank, tasks, ctl, cbases, ntasks, cand, lr);
}

// ---------------------------------------------------------------------------
// Medium tasks (short of sampling, too long for k_small), one block each,
// beside the long-task passes: the first-digit histogram of a read of the
// task in LDS (in the space of v, cap >= 2048 entries), the selection, then
// finish_task re-reads the task while it is still cached.  Results sorted
// here (m <= cap).
// ---------------------------------------------------------------------------
template <int EMAX>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MED_WAVES_PER_EU))) void k_medium(const TaskDesc* __restrict__ tasks,
                                                  uint64_t* __restrict__ cand, uint32_t cap) {
  extern __shared__ uint64_t v[];
  __shared__ FinShared S;
  const TaskDesc d = tasks[blockIdx.x];
  uint32_t* lh = reinterpret_cast<uint32_t*>(v);
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) lh[i] = 0;
  __syncthreads();
  const float sc = task_scaler(d);
  const bool scale = task_scales(d);
  const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
  stream_range<FIN_UNR>(d.in, 0, d.len,
                        [&](uint32_t x, uint32_t) { atomicAdd(&lh[rank_key(x, sc, xm, scale) >> 20], 1u); });
  __syncthreads();
  uint32_t h[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) h[j] = lh[threadIdx.x * 16 + j];
  select_first(d, h, NO_BIN, 0u, cap, &S.ctl, S.wsum, S.res);
  __syncthreads();  // the selection is published and the histogram read: v is free
  const TaskCtl c = S.ctl;
  finish_task<EMAX>(d, c, cand, v, cap, INT64_MAX, 1, false, S);
}

// ---------------------------------------------------------------------------
// Short tasks: one block does the whole top-k in LDS.  Dynamic LDS holds
// 2 * cap composites (the task, then the chosen ones).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_small(const TaskDesc* __restrict__ tasks, uint32_t cap) {
  extern __shared__ uint64_t sm[];
  __shared__ uint32_t lh[HIST_BINS];
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  __shared__ uint32_t ectr;
  __shared__ unsigned long long vb[2];
  const TaskDesc d = tasks[blockIdx.x];
  const uint32_t n = (uint32_t)d.len;
  const uint32_t m = d.k < (int32_t)n ? (uint32_t)d.k : n;
  uint64_t* __restrict__ all = sm;
  uint64_t* __restrict__ sel = sm + cap;
  if (n > cap) return;  // host routes only short tasks here
  const float sc = task_scaler(d);
  const bool scale = task_scales(d);
  const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
  if (threadIdx.x == 0) {
    ectr = 0;
    vb[0] = ~0ull;
    vb[1] = 0ull;
  }
  __syncthreads();
  uint64_t a = ~0ull, o = 0ull;
  for (uint32_t b = threadIdx.x; b < n; b += BLOCK * LOADS_IN_FLIGHT) {
    uint32_t xs[LOADS_IN_FLIGHT];
#pragma unroll
    for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
      const uint32_t i = b + (uint32_t)u * BLOCK;
      xs[u] = (i < n) ? load_elem(d.in, i) : 0u;
    }
#pragma unroll
    for (int u = 0; u < LOADS_IN_FLIGHT; ++u) {
      const uint32_t i = b + (uint32_t)u * BLOCK;
      if (i < n) {
        const uint64_t x = ((uint64_t)rank_key(xs[u], sc, xm, scale) << 32) | i;
        all[i] = x;
        a &= x;
        o |= x;
      }
    }
  }
  block_and_or(a, o, vb);  // also publishes `all`
  int s = 64;
  uint64_t P = 0;
  if (m > 0 && m < n)
    s = select_digits(LdsSrc{all}, n, 64, P, m, (uint64_t)(vb[0] ^ vb[1]), (uint64_t)vb[0], lh, wsum, res);
  if (m > 0) {
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
      const uint64_t x = all[i];
      const bool take = (m >= n) ? true : ((x >> s) <= P);
      const uint64_t tm = ballot(take);
      if (tm) {
        const uint32_t slot = wave_append(take, tm, &ectr);
        if (take && slot < m) {
          const uint32_t pos = (uint32_t)x;
          const uint32_t v = scale ? load_elem(d.in, pos) : ekey_inv((uint32_t)(x >> 32) ^ xm);
          sel[slot] = out_composite(d, v, pos);
        }
      }
    }
  }
  __syncthreads();
  bitonic_sort_asc(sel, m);
  write_sorted(d, (const uint64_t*)sel, m);
}

// ---------------------------------------------------------------------------
// Diagnostics: wave and block scans as the selection kernels use them.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_diag_scan(const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ ow,
                                                     uint32_t* __restrict__ ob, int64_t n) {
  __shared__ uint32_t wsum[BLOCK / 64];
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t v = (i < n) ? in[i] : 0u;
  const uint32_t w = wave_incl_scan(v);
  const uint32_t b = block_incl_scan(v, wsum);
  if (i < n) {
    ow[i] = w;
    ob[i] = b;
  }
}

// Plain streaming copy (16-byte loads and stores), the reference point for
// HBM traffic counters: it reads and writes exactly `n16` * 16 bytes.
__global__ __launch_bounds__(BLOCK) void k_diag_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst,
                                                     int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n16; i += (int64_t)gridDim.x * BLOCK)
    glb(dst)[i] = __builtin_nontemporal_load(glb(src) + i);
}

}  // namespace RADIK_KNS
}  // namespace radik
