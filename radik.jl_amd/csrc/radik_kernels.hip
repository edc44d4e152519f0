// radik_kernels.hip -- gfx950 kernels of the radix top-k path.
//
// Selection works on the 64-bit composite  C = rank_key(x) << 32 | position,
// smaller is better, so the top-k of a task are its k smallest composites and
// ties at the k-th value fall to the lowest position.  The first digit
// (rank-key bits [31:20], the reference's pass-1 digit, count_bin.jl:78-203,
// algorithm.jl:221-224) is resolved over the raw input by two streaming passes:
//
//   k_hist     4096-bin histogram per task (LDS ds_add, one global atomic per
//              non-zero bin per block)                       ~ count_bin_ex_kernel!
//   k_select   block scan of the 4096 bins -> winner bound, candidate bin,
//              residual k                                    ~ select_bin_kernel!
//   k_compact  second read: winners (digit < bin) go straight to the output
//              slots, candidates (digit == bin) are spilled as composites;
//              wave ballot + one atomic per wave              ~ select_candidate_ex_kernel!
//   k_refine   per task: the remaining digits (rank bits [19:8], [7:0] as in
//              algorithm.jl:247-299, then position digits for exact ties) over
//              the spilled candidates, in LDS when they fit  ~ passes 2/3 + filter
//   k_sort     per task: bitonic sort of the k results by (output key,
//              position), indices gathered                    ~ bitonic_sort!/merge_sort
//   k_small    one block per short task: everything above in LDS, one launch.
//
// Streaming kernels (k_hist, k_compact) split the launch's concatenated element
// space into equal chunks, one per block, so uneven task lengths balance at
// element granularity; a chunk that crosses task boundaries flushes per task.

#include "radik_device.h"

namespace radik {

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

// Visits elements [s, e) of task array p: scalar head up to 16-byte alignment,
// UNR float4 loads in flight per thread over the body, scalar tail.  f(x, pos).
template <int UNR, typename F>
__device__ __forceinline__ void stream_range(const float* __restrict__ p, int64_t s, int64_t e, F&& f) {
  const int64_t tid = threadIdx.x;
  int64_t a = s;
  // elements until p + a is 16-byte aligned
  int64_t mis = (int64_t)(((uintptr_t)(p + a) >> 2) & 3u);
  int64_t head_end = a + ((4 - mis) & 3);
  if (head_end > e) head_end = e;
  for (int64_t i = a + tid; i < head_end; i += BLOCK) f(p[i], (uint32_t)i);
  a = head_end;
  const int64_t nvec = (e - a) >> 2;
  const f32x4* __restrict__ pv = reinterpret_cast<const f32x4*>(p + a);
  int64_t v = tid;
  for (; v + (int64_t)(UNR - 1) * BLOCK < nvec; v += (int64_t)UNR * BLOCK) {
    f32x4 r[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) r[u] = __builtin_nontemporal_load(pv + v + (int64_t)u * BLOCK);
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const uint32_t q = (uint32_t)(a + ((v + (int64_t)u * BLOCK) << 2));
      f(r[u].x, q);
      f(r[u].y, q + 1);
      f(r[u].z, q + 2);
      f(r[u].w, q + 3);
    }
  }
  for (; v < nvec; v += BLOCK) {
    f32x4 r = __builtin_nontemporal_load(pv + v);
    const uint32_t q = (uint32_t)(a + (v << 2));
    f(r.x, q);
    f(r.y, q + 1);
    f(r.z, q + 2);
    f(r.w, q + 3);
  }
  for (int64_t i = a + (nvec << 2) + tid; i < e; i += BLOCK) f(p[i], (uint32_t)i);
}

// ---------------------------------------------------------------------------
// Pass A: first-digit histogram.
// ---------------------------------------------------------------------------
template <int UNR>
__global__ __launch_bounds__(BLOCK) void k_hist(const TaskDesc* __restrict__ tasks,
                                                const int64_t* __restrict__ bases, int ntasks,
                                                int64_t total, int64_t chunk,
                                                uint32_t* __restrict__ hist) {
  __shared__ uint32_t lh[HIST_BINS];
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) lh[i] = 0;
  int64_t g0 = (int64_t)blockIdx.x * chunk;
  const int64_t g1 = min(g0 + chunk, total);
  if (g0 >= g1) return;
  int t = find_task(bases, ntasks, g0);
  __syncthreads();
  while (true) {
    const TaskDesc d = tasks[t];
    const int64_t tb = bases[t];
    const int64_t te = tb + d.len;
    const int64_t se = min(g1, te);
    const float sc = task_scaler(d);
    const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
    stream_range<UNR>(d.in, g0 - tb, se - tb, [&](float x, uint32_t) {
      atomicAdd(&lh[rank_key(x, sc, xm) >> 20], 1u);
    });
    __syncthreads();
    uint32_t* __restrict__ h = hist + (size_t)t * HIST_BINS;
    for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) {
      const uint32_t c = lh[i];
      if (c) {
        atomicAdd(&h[i], c);
        lh[i] = 0;
      }
    }
    if (se >= g1) break;
    g0 = se;
    ++t;
    while (t + 1 < ntasks && bases[t + 1] <= g0) ++t;  // skip empty tasks
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// First-digit selection, one block per task.  Reads and clears the task's
// histogram so the buffer is zero again for the next launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_select(const TaskDesc* __restrict__ tasks,
                                                  uint32_t* __restrict__ hist,
                                                  TaskCtl* __restrict__ ctl) {
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  const int t = blockIdx.x;
  const TaskDesc d = tasks[t];
  uint4* __restrict__ h4 = reinterpret_cast<uint4*>(hist + (size_t)t * HIST_BINS) + threadIdx.x * 4;
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
  const int64_t len = d.len;
  int64_t m = d.k < len ? (int64_t)d.k : len;
  if (threadIdx.x == 0) res[0] = NO_BIN;  // published by the barriers inside the scan
  if (m > 0 && m < len) block_find_bin<16>(c, (uint32_t)m, wsum, res);
  if (threadIdx.x == 0) {
    TaskCtl o = {};
    o.b_cand = NO_BIN;
    o.krem = 0;
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
    o.cand_ctr = 0;
    ctl[t] = o;
  }
}

// Stores a result (value, position) into output slot `slot` of task d.
__device__ __forceinline__ void put_result(const TaskDesc& d, uint32_t slot, float v, uint32_t pos) {
  d.val_out[slot] = v;
  if (d.flags & F_IDX64) reinterpret_cast<int64_t*>(d.idx_out)[slot] = (int64_t)pos;
  else reinterpret_cast<int32_t*>(d.idx_out)[slot] = (int32_t)pos;
}

// Wave-aggregated append: every lane with `p` gets a distinct slot from
// *ctr; one atomic per wave.  Must be reached by all active lanes together.
__device__ __forceinline__ uint32_t wave_append(bool p, uint64_t mask, uint32_t* ctr) {
  const uint64_t act = ballot(true);
  uint32_t base = 0;
  if (lanes_below(act) == 0) base = atomicAdd(ctr, (uint32_t)__builtin_popcountll(mask));
  base = __builtin_amdgcn_readfirstlane(base);
  return base + lanes_below(mask);
}

// ---------------------------------------------------------------------------
// Pass B: compaction.  Winners go to the output slots [0, nwin) as staging
// (value, position); candidates to the task's region of the spill buffer.
// ---------------------------------------------------------------------------
template <int UNR>
__global__ __launch_bounds__(BLOCK) void k_compact(const TaskDesc* __restrict__ tasks,
                                                   const int64_t* __restrict__ bases, int ntasks,
                                                   int64_t total, int64_t chunk,
                                                   TaskCtl* __restrict__ ctl,
                                                   uint64_t* __restrict__ cand) {
  int64_t g0 = (int64_t)blockIdx.x * chunk;
  const int64_t g1 = min(g0 + chunk, total);
  if (g0 >= g1) return;
  int t = find_task(bases, ntasks, g0);
  while (true) {
    const TaskDesc d = tasks[t];
    const int64_t tb = bases[t];
    const int64_t te = tb + d.len;
    const int64_t se = min(g1, te);
    const uint32_t bw = ctl[t].b_win;
    const uint32_t bc = ctl[t].b_cand;
    const uint32_t nwin = ctl[t].nwin;
    if (bw > 0 || bc != NO_BIN) {
      const float sc = task_scaler(d);
      const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
      uint32_t* wctr = &ctl[t].win_ctr;
      uint32_t* cctr = &ctl[t].cand_ctr;
      uint64_t* __restrict__ cb = cand + d.base;
      const uint32_t clen = (uint32_t)d.len;
      stream_range<UNR>(d.in, g0 - tb, se - tb, [&](float x, uint32_t pos) {
        const uint32_t rk = rank_key(x, sc, xm);
        const uint32_t dg = rk >> 20;
        const bool w = dg < bw;
        const bool c = dg == bc;
        const uint64_t mw = ballot(w);
        const uint64_t mc = ballot(c);
        if (mw) {
          const uint32_t slot = wave_append(w, mw, wctr);
          if (w && slot < nwin) put_result(d, slot, x, pos);
        }
        if (mc) {
          const uint32_t slot = wave_append(c, mc, cctr);
          if (c && slot < clen) cb[slot] = ((uint64_t)rk << 32) | pos;
        }
      });
    }
    if (se >= g1) break;
    g0 = se;
    ++t;
    while (t + 1 < ntasks && bases[t + 1] <= g0) ++t;
  }
}

// ---------------------------------------------------------------------------
// Digit selection over composites (shared by k_refine and k_small).
// Digits below the first: rank bits [19:8], [7:0], then position bits
// [31:20], [19:8], [7:0].  Starting from prefix P (the composite bits above
// `top` already fixed) and need = how many of the prefix-matching composites
// to take, finds the shortest prefix at which the bin holding the need-th
// smallest is taken whole.  Returns the shift s: the chosen composites are
// exactly those with (x >> s) <= P (P updated), among those matching the
// initial prefix.
// ---------------------------------------------------------------------------
template <typename Src>
__device__ __forceinline__ int select_digits(Src src, uint32_t n, int top, uint64_t& P, uint32_t need,
                                             uint32_t* lh, uint32_t* wsum, uint32_t* res) {
  const int shifts[6] = {52, 40, 32, 20, 8, 0};
  const int widths[6] = {12, 12, 8, 12, 12, 8};
  int s_final = top;
  for (int q = 0; q < 6; ++q) {
    const int s = shifts[q];
    const int w = widths[q];
    if (s + w > top) continue;  // digit already fixed
    const uint32_t nb = 1u << w;
    for (uint32_t i = threadIdx.x; i < nb; i += BLOCK) lh[i] = 0;
    __syncthreads();
    const int hs = s + w;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
      const uint64_t x = src[i];
      const bool match = (hs >= 64) ? true : ((x >> hs) == P);
      if (match) atomicAdd(&lh[(uint32_t)(x >> s) & (nb - 1)], 1u);
    }
    __syncthreads();
    if (w == 12) {
      uint32_t c[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) c[j] = lh[threadIdx.x * 16 + j];
      block_find_bin<16>(c, need, wsum, res);
    } else {
      uint32_t c[1];
      c[0] = lh[threadIdx.x];
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
// Refinement of the spilled candidates, one block per task.
// ---------------------------------------------------------------------------
struct LdsSrc {
  const uint64_t* p;
  __device__ uint64_t operator[](uint32_t i) const { return p[i]; }
};
struct GlobalSrc {
  const uint64_t* __restrict__ p;
  __device__ uint64_t operator[](uint32_t i) const { return p[i]; }
};

template <typename Src>
__device__ __forceinline__ void refine_emit(const TaskDesc& d, const TaskCtl& c, Src src, uint32_t n,
                                            uint32_t* lh, uint32_t* wsum, uint32_t* res, uint32_t* ectr) {
  uint64_t P = c.b_cand;
  const int s = select_digits(src, n, 52, P, c.krem, lh, wsum, res);
  const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
  const bool scale = (d.flags & F_WITHSCALE) != 0;
  const uint32_t lim = c.m - c.nwin;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
    const uint64_t x = src[i];
    if ((x >> s) <= P) {
      const uint32_t slot = atomicAdd(ectr, 1u);
      if (slot < lim) {
        uint32_t pos = (uint32_t)x;
        if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
        const float v = scale ? d.in[pos] : fkey_inv((uint32_t)(x >> 32) ^ xm);
        put_result(d, c.nwin + slot, v, pos);
      }
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_refine(const TaskDesc* __restrict__ tasks,
                                                  const TaskCtl* __restrict__ ctl,
                                                  const uint64_t* __restrict__ cand,
                                                  uint32_t lds_cap) {
  extern __shared__ uint64_t cbuf[];
  __shared__ uint32_t lh[HIST_BINS];
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  __shared__ uint32_t ectr;
  const int t = blockIdx.x;
  const TaskCtl c = ctl[t];
  if (c.krem == 0) return;
  const TaskDesc d = tasks[t];
  uint32_t n = c.cand_ctr;
  if (n > (uint32_t)d.len) n = (uint32_t)d.len;
  if (threadIdx.x == 0) ectr = 0;
  const uint64_t* __restrict__ src = cand + d.base;
  if (n <= lds_cap) {
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) cbuf[i] = src[i];
    __syncthreads();
    refine_emit(d, c, LdsSrc{cbuf}, n, lh, wsum, res, &ectr);
  } else {
    __syncthreads();
    refine_emit(d, c, GlobalSrc{src}, n, lh, wsum, res, &ectr);
  }
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

// Final order key of a result: (ASCEND ? key : ~key) << 32 | position.
__device__ __forceinline__ uint64_t out_composite(const TaskDesc& d, float v, uint32_t pos) {
  const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
  return ((uint64_t)(fkey(v) ^ am) << 32) | pos;
}

// Writes the k output slots of task d from sorted composites (padding past m).
template <typename Src>
__device__ __forceinline__ void write_sorted(const TaskDesc& d, Src s, uint32_t m) {
  const uint32_t am = (d.flags & F_ASCEND) ? 0u : 0xFFFFFFFFu;
  const bool withidx = (d.flags & F_WITHIDXIN) != 0;
  const bool i64 = (d.flags & F_IDX64) != 0;
  for (uint32_t i = threadIdx.x; i < (uint32_t)d.k; i += BLOCK) {
    float v = 0.0f;
    int64_t id = -1;
    if (i < m) {
      const uint64_t x = s[i];
      uint32_t pos = (uint32_t)x;
      if ((int64_t)pos >= d.len) pos = (uint32_t)(d.len - 1);  // never read outside the task
      v = fkey_inv((uint32_t)(x >> 32) ^ am);
      if (withidx) id = i64 ? reinterpret_cast<const int64_t*>(d.idx_in)[pos]
                            : (int64_t)reinterpret_cast<const int32_t*>(d.idx_in)[pos];
      else id = (int64_t)pos + 1;
    }
    d.val_out[i] = v;
    if (i64) reinterpret_cast<int64_t*>(d.idx_out)[i] = id;
    else reinterpret_cast<int32_t*>(d.idx_out)[i] = (int32_t)id;
  }
}

// ---------------------------------------------------------------------------
// Final sort of the staged results, one block per task.  LDS when m fits,
// else in the task's (now free) spill region, which holds len >= m slots.
// ---------------------------------------------------------------------------
struct GlobalPtr {
  uint64_t* __restrict__ p;
  __device__ uint64_t& operator[](uint32_t i) const { return p[i]; }
};

__global__ __launch_bounds__(BLOCK) void k_sort(const TaskDesc* __restrict__ tasks,
                                                const TaskCtl* __restrict__ ctl,
                                                uint64_t* __restrict__ cand, uint32_t lds_cap) {
  extern __shared__ uint64_t sb[];
  const int t = blockIdx.x;
  const TaskDesc d = tasks[t];
  const uint32_t m = ctl[t].m;
  const bool i64 = (d.flags & F_IDX64) != 0;
  if (m <= lds_cap) {
    for (uint32_t i = threadIdx.x; i < m; i += BLOCK) {
      const uint32_t pos = i64 ? (uint32_t)reinterpret_cast<const int64_t*>(d.idx_out)[i]
                               : (uint32_t)reinterpret_cast<const int32_t*>(d.idx_out)[i];
      sb[i] = out_composite(d, d.val_out[i], pos);
    }
    __syncthreads();
    bitonic_sort_asc(sb, m);
    write_sorted(d, (const uint64_t*)sb, m);
  } else {
    uint64_t* __restrict__ g = cand + d.base;
    for (uint32_t i = threadIdx.x; i < m; i += BLOCK) {
      const uint32_t pos = i64 ? (uint32_t)reinterpret_cast<const int64_t*>(d.idx_out)[i]
                               : (uint32_t)reinterpret_cast<const int32_t*>(d.idx_out)[i];
      g[i] = out_composite(d, d.val_out[i], pos);
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
// Short tasks: one block does the whole top-k in LDS.  Dynamic LDS holds
// 2 * cap composites (the task, then the chosen ones).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_small(const TaskDesc* __restrict__ tasks, uint32_t cap) {
  extern __shared__ uint64_t sm[];
  __shared__ uint32_t lh[HIST_BINS];
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  __shared__ uint32_t ectr;
  const TaskDesc d = tasks[blockIdx.x];
  const uint32_t n = (uint32_t)d.len;
  const uint32_t m = d.k < (int32_t)n ? (uint32_t)d.k : n;
  uint64_t* __restrict__ all = sm;
  uint64_t* __restrict__ sel = sm + cap;
  if (n > cap) return;  // host routes only short tasks here
  const float sc = task_scaler(d);
  const uint32_t xm = (d.flags & F_LARGEST) ? 0xFFFFFFFFu : 0u;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK)
    all[i] = ((uint64_t)rank_key(d.in[i], sc, xm) << 32) | i;
  if (threadIdx.x == 0) ectr = 0;
  __syncthreads();
  int s = 64;
  uint64_t P = 0;
  if (m > 0 && m < n) s = select_digits(LdsSrc{all}, n, 64, P, m, lh, wsum, res);
  const bool scale = (d.flags & F_WITHSCALE) != 0;
  if (m > 0) {
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
      const uint64_t x = all[i];
      const bool take = (m >= n) ? true : ((x >> s) <= P);
      if (take) {
        const uint32_t slot = atomicAdd(&ectr, 1u);
        if (slot < m) {
          const uint32_t pos = (uint32_t)x;
          const float v = scale ? d.in[pos] : fkey_inv((uint32_t)(x >> 32) ^ xm);
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

// Explicit instantiations used by the host.
template __global__ void k_hist<1>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, uint32_t*);
template __global__ void k_hist<2>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, uint32_t*);
template __global__ void k_hist<4>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, uint32_t*);
template __global__ void k_compact<1>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, TaskCtl*, uint64_t*);
template __global__ void k_compact<2>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, TaskCtl*, uint64_t*);
template __global__ void k_compact<4>(const TaskDesc*, const int64_t*, int, int64_t, int64_t, TaskCtl*, uint64_t*);

}  // namespace radik
