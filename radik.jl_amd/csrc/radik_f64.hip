// radik_f64.hip -- top-k over Float64 keys.
//
// The main path ranks elements by a 32-bit key inside 64-bit composites
// (key << 32 | position); a Float64 key needs all 64 bits, so Float64 tasks
// take this exact path of their own, with the same contract: the m =
// min(k, len) best by (rank, position), ties to the lowest position, sorted
// by (ASCEND ? key : ~key, position).
//
//   k64_init     per task: the state (nothing to select when m = 0 or m >= len)
//   k64_hist     per chunk (a task's elements in chunks of F64_CHUNK): the
//                histogram of the next rank digit over the elements whose
//                rank carries the prefix fixed so far (digits of 12, 12, 12,
//                12, 12 and 4 bits from the top)
//   k64_select   per task: the digit holding the residual need-th element;
//                the prefix grows by it, or every element of that digit is
//                in and the selection ends
//   k64_ties     per chunk: elements equal to the full 64-bit threshold
//   k64_collect  per chunk: the chosen elements (rank prefix below the
//                threshold, or equal and among the first `need` equal ones
//                by position: a per-chunk base plus an ordered block scan),
//                appended to the output slots as (value, position)
//   k64_sort     per task: the m results sorted in LDS by (output key,
//                position), indices mapped, padding past m
// Each element is read once per digit pass (6), once for the tie count and
// once for the collection; Float64 tasks take k <= F64_MAX_K.
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "../../include/radik.h"
#include "radik_host.h"

namespace radik {

constexpr int64_t F64_CHUNK = 1 << 16;  // elements per chunk block
constexpr int F64_PASSES = 6;
__device__ constexpr int kF64Width[F64_PASSES] = {12, 12, 12, 12, 12, 4};
__device__ constexpr int kF64Shift[F64_PASSES] = {52, 40, 28, 16, 4, 0};

struct F64Task {
  const uint64_t* in;
  const void* idx_in;
  uint64_t* val_out;
  void* idx_out;
  int64_t len;
  int64_t c0;  // first chunk
  int32_t k;
  uint32_t flags;
  int64_t pad;
};
static_assert(sizeof(F64Task) == 64, "F64Task layout");

struct F64Chunk {
  int64_t s, e;  // element range in the task
  uint32_t t;
  uint32_t pad;
  int64_t pad2;
};

struct F64Ctl {
  uint64_t P;       // rank prefix fixed so far: rank >> shift == P selects the threshold group
  uint32_t shift;   // 64 before the first pass
  uint32_t need;    // elements still to take inside the threshold group
  uint32_t active;  // more digits to resolve
  uint32_t mode;    // 0: nothing, 1: take every element, 2: take rank >> shift < P, and == P all,
                    // 3: take rank < P, and the first `need` equal to P by position
  uint32_t m;
  uint32_t ctr;     // output slots appended
};

__device__ __forceinline__ uint64_t f64_rank(uint64_t bits, uint32_t flags) {
  const uint64_t key = bits ^ ((bits >> 63) ? ~0ull : 0x8000000000000000ull);
  return (flags & F_LARGEST) ? ~key : key;
}

__device__ __forceinline__ bool f64_match(uint64_t r, const F64Ctl& c) {
  return c.shift >= 64 ? true : (r >> c.shift) == c.P;
}

__global__ __launch_bounds__(BLOCK) void k64_init(const F64Task* __restrict__ tasks, F64Ctl* __restrict__ ctl,
                                                  int ntasks) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= ntasks) return;
  const F64Task d = tasks[t];
  const int64_t m = d.k < d.len ? (int64_t)d.k : d.len;
  F64Ctl c = {};
  c.shift = 64;
  c.m = (uint32_t)m;
  c.need = (uint32_t)m;
  if (m <= 0) c.mode = 0;
  else if (m >= d.len) c.mode = 1;
  else c.active = 1;
  ctl[t] = c;
}

__global__ __launch_bounds__(BLOCK) void k64_hist(const F64Task* __restrict__ tasks, const F64Chunk* __restrict__ chunks,
                                                  const F64Ctl* __restrict__ ctl, uint32_t* __restrict__ hist,
                                                  int pass) {
  __shared__ uint32_t lh[HIST_BINS];
  const F64Chunk ch = chunks[blockIdx.x];
  const F64Ctl c = ctl[ch.t];
  if (!c.active) return;
  const F64Task d = tasks[ch.t];
  const int ds = kF64Shift[pass];
  const uint32_t mask = (1u << kF64Width[pass]) - 1u;
  for (int i = threadIdx.x; i < HIST_BINS; i += BLOCK) lh[i] = 0;
  __syncthreads();
  for (int64_t i = ch.s + threadIdx.x; i < ch.e; i += BLOCK) {
    const uint64_t r = f64_rank(d.in[i], d.flags);
    if (f64_match(r, c)) atomicAdd(&lh[(uint32_t)(r >> ds) & mask], 1u);
  }
  __syncthreads();
  uint32_t* h = hist + (size_t)ch.t * HIST_BINS;
  for (int i = threadIdx.x; i <= (int)mask; i += BLOCK)
    if (lh[i]) atomicAdd(&h[i], lh[i]);
}

__global__ __launch_bounds__(BLOCK) void k64_select(F64Ctl* __restrict__ ctl, uint32_t* __restrict__ hist, int pass) {
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t res[3];
  const int t = blockIdx.x;
  F64Ctl c = ctl[t];
  if (!c.active) return;
  uint32_t* h = hist + (size_t)t * HIST_BINS;
  uint32_t b16[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    b16[j] = h[threadIdx.x * 16 + j];
    h[threadIdx.x * 16 + j] = 0u;  // zero again for the next pass and launch
  }
  if (threadIdx.x == 0) res[0] = NO_BIN;
  block_find_bin<16>(b16, c.need, wsum, res);
  if (threadIdx.x != 0) return;
  if (res[0] == NO_BIN) {  // (cannot happen: the histogram holds need or more)
    c.active = 0;
    c.mode = 0;
    ctl[t] = c;
    return;
  }
  const uint32_t b = res[0], before = res[1], cnt = res[2];
  const int w = kF64Width[pass];
  c.P = (c.shift >= 64 ? 0ull : (c.P << w)) | b;
  c.shift = (uint32_t)kF64Shift[pass];
  c.need -= before;
  if (c.need == cnt) {
    c.active = 0;
    c.mode = 2;  // the whole group is in
  } else if (c.shift == 0) {
    c.active = 0;
    c.mode = 3;  // exact ties at the threshold: the first `need` by position
  }
  ctl[t] = c;
}

__global__ __launch_bounds__(BLOCK) void k64_ties(const F64Task* __restrict__ tasks, const F64Chunk* __restrict__ chunks,
                                                  const F64Ctl* __restrict__ ctl, uint32_t* __restrict__ ccnt) {
  __shared__ uint32_t n;
  const F64Chunk ch = chunks[blockIdx.x];
  const F64Ctl c = ctl[ch.t];
  if (threadIdx.x == 0) n = 0;
  __syncthreads();
  if (c.mode == 3) {
    const F64Task d = tasks[ch.t];
    uint32_t mine = 0;
    for (int64_t i = ch.s + threadIdx.x; i < ch.e; i += BLOCK) mine += f64_rank(d.in[i], d.flags) == c.P ? 1u : 0u;
    if (mine) atomicAdd(&n, mine);
  }
  __syncthreads();
  if (threadIdx.x == 0) ccnt[blockIdx.x] = n;
}

__device__ __forceinline__ void f64_put(const F64Task& d, uint32_t slot, uint64_t bits, uint32_t pos) {
  d.val_out[slot] = bits;
  if (d.flags & F_IDX64) reinterpret_cast<int64_t*>(d.idx_out)[slot] = (int64_t)pos;
  else reinterpret_cast<int32_t*>(d.idx_out)[slot] = (int32_t)pos;
}

__global__ __launch_bounds__(BLOCK) void k64_collect(const F64Task* __restrict__ tasks,
                                                     const F64Chunk* __restrict__ chunks, F64Ctl* __restrict__ ctl,
                                                     const uint32_t* __restrict__ ccnt) {
  __shared__ uint32_t wsum[BLOCK / 64];
  __shared__ uint32_t tot;
  const F64Chunk ch = chunks[blockIdx.x];
  const F64Ctl c = ctl[ch.t];
  if (c.mode == 0) return;
  const F64Task d = tasks[ch.t];
  uint32_t* ctr = &ctl[ch.t].ctr;
  if (c.mode != 3) {  // no order needed: every chosen element is appended as met
    for (int64_t i = ch.s + threadIdx.x; i - threadIdx.x < ch.e; i += BLOCK) {
      bool take = false;
      uint64_t bits = 0;
      if (i < ch.e) {
        bits = d.in[i];
        const uint64_t r = f64_rank(bits, d.flags);
        take = c.mode == 1 || (r >> c.shift) <= c.P;
      }
      const uint64_t mk = ballot(take);
      if (mk) {
        const uint32_t slot = wave_append(take, mk, ctr);
        if (take && slot < c.m) f64_put(d, slot, bits, (uint32_t)i);
      }
    }
    return;
  }
  // exact ties: the threshold-equal elements before this chunk, then an
  // ordered count over the chunk's tiles
  uint32_t base = 0;
  for (int64_t j = d.c0 + threadIdx.x; j < (int64_t)blockIdx.x; j += BLOCK) base += ccnt[j];
  base = block_incl_scan(base, wsum);  // (thread BLOCK - 1 holds the total)
  if (threadIdx.x == BLOCK - 1) tot = base;
  __syncthreads();
  base = tot;
  __syncthreads();
  for (int64_t t0 = ch.s; t0 < ch.e; t0 += BLOCK) {
    const int64_t i = t0 + threadIdx.x;
    uint64_t bits = 0, r = ~0ull;
    if (i < ch.e) {
      bits = d.in[i];
      r = f64_rank(bits, d.flags);
    }
    const uint32_t eq = (i < ch.e && r == c.P) ? 1u : 0u;
    const uint32_t incl = block_incl_scan(eq, wsum);
    if (threadIdx.x == BLOCK - 1) tot = incl;
    const bool take = i < ch.e && (r < c.P || (eq && base + incl - 1 < c.need));
    const uint64_t mk = ballot(take);
    if (mk) {
      const uint32_t slot = wave_append(take, mk, ctr);
      if (take && slot < c.m) f64_put(d, slot, bits, (uint32_t)i);
    }
    __syncthreads();
    base += tot;
    __syncthreads();
  }
}

// Sorts the m (value, position) results of a task by (output key, position)
// in LDS: keys (8 bytes) and positions (4 bytes) in two arrays of p2 >= m.
__global__ __launch_bounds__(BLOCK) void k64_sort(const F64Task* __restrict__ tasks, const F64Ctl* __restrict__ ctl,
                                                  uint32_t p2max) {
  extern __shared__ uint64_t smem[];
  uint64_t* key = smem;
  uint32_t* pos = reinterpret_cast<uint32_t*>(smem + p2max);
  const int t = blockIdx.x;
  const F64Task d = tasks[t];
  const uint32_t m = ctl[t].m;
  const uint64_t am = (d.flags & F_ASCEND) ? 0ull : ~0ull;
  const bool i64 = (d.flags & F_IDX64) != 0;
  uint32_t p2 = 1;
  while (p2 < m) p2 <<= 1;
  for (uint32_t i = threadIdx.x; i < p2; i += BLOCK) {
    if (i < m) {
      const uint64_t bits = d.val_out[i];
      key[i] = (bits ^ ((bits >> 63) ? ~0ull : 0x8000000000000000ull)) ^ am;
      pos[i] = i64 ? (uint32_t)reinterpret_cast<const int64_t*>(d.idx_out)[i]
                   : (uint32_t)reinterpret_cast<const int32_t*>(d.idx_out)[i];
    } else {
      key[i] = ~0ull;
      pos[i] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  for (uint32_t size = 2; size <= p2; size <<= 1) {
    for (uint32_t j = size >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < p2; i += BLOCK) {
        const uint32_t q = i ^ j;
        if (q > i) {
          const bool up = (i & size) == 0;
          const uint64_t ka = key[i], kb = key[q];
          const uint32_t pa = pos[i], pb = pos[q];
          const bool gt = ka > kb || (ka == kb && pa > pb);
          if (gt == up) {
            key[i] = kb;
            key[q] = ka;
            pos[i] = pb;
            pos[q] = pa;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = threadIdx.x; i < (uint32_t)d.k; i += BLOCK) {
    uint64_t v = 0;
    int64_t id = -1;
    if (i < m) {
      const uint64_t k0 = key[i] ^ am;
      v = (k0 >> 63) ? (k0 ^ 0x8000000000000000ull) : ~k0;
      const uint32_t p = pos[i];
      if (d.flags & F_WITHIDXIN)
        id = i64 ? reinterpret_cast<const int64_t*>(d.idx_in)[p] : (int64_t)reinterpret_cast<const int32_t*>(d.idx_in)[p];
      else
        id = (int64_t)p + 1;
    }
    d.val_out[i] = v;
    if (i64) reinterpret_cast<int64_t*>(d.idx_out)[i] = id;
    else reinterpret_cast<int32_t*>(d.idx_out)[i] = (int32_t)id;
  }
}

// ---------------------------------------------------------------------------
// Host side: buffers of the Float64 path of one pipeline (grow-only).
// ---------------------------------------------------------------------------
struct F64Pipe {
  std::vector<F64Task> tasks;
  std::vector<F64Chunk> chunks;
  F64Task* d_tasks = nullptr;
  F64Chunk* d_chunks = nullptr;
  F64Ctl* d_ctl = nullptr;
  uint32_t* d_hist = nullptr;
  uint32_t* d_ccnt = nullptr;
  size_t cap_t = 0, cap_c = 0;
  uint32_t p2max = 1;
  ~F64Pipe() {
    if (d_tasks) (void)hipFree(d_tasks);
    if (d_chunks) (void)hipFree(d_chunks);
    if (d_ctl) (void)hipFree(d_ctl);
    if (d_hist) (void)hipFree(d_hist);
    if (d_ccnt) (void)hipFree(d_ccnt);
  }
};

void f64_destroy(F64Pipe* p) { delete p; }

int f64_stage(F64Pipe** pp, const TaskSpec* specs, int64_t n, hipStream_t s) {
  if (!*pp) *pp = new F64Pipe();
  F64Pipe& p = **pp;
  p.tasks.resize((size_t)n);
  p.chunks.clear();
  uint32_t maxm = 1;
  for (int64_t t = 0; t < n; ++t) {
    const TaskSpec& sp = specs[t];
    if (sp.k > RADIK_F64_MAX_K)
      return set_error(RADIK_EINVAL, "Float64 keys take k <= %d, got %lld", RADIK_F64_MAX_K, (long long)sp.k);
    F64Task& d = p.tasks[(size_t)t];
    d.in = (const uint64_t*)sp.in;
    d.idx_in = sp.idx_in;
    d.val_out = (uint64_t*)sp.val_out;
    d.idx_out = sp.idx_out;
    d.len = sp.len;
    d.k = (int32_t)sp.k;
    d.flags = sp.flags;
    d.pad = 0;
    d.c0 = (int64_t)p.chunks.size();
    for (int64_t a = 0; a < sp.len; a += F64_CHUNK) {
      F64Chunk c;
      c.s = a;
      c.e = std::min(sp.len, a + F64_CHUNK);
      c.t = (uint32_t)t;
      c.pad = 0;
      c.pad2 = 0;
      p.chunks.push_back(c);
    }
    const int64_t m = sp.k < sp.len ? sp.k : sp.len;
    if ((uint32_t)m > maxm) maxm = (uint32_t)m;
  }
  uint32_t p2 = 1;
  while (p2 < maxm) p2 <<= 1;
  p.p2max = p2;
  if ((size_t)n > p.cap_t) {
    RADIK_HIP(hipStreamSynchronize(s));
    size_t c = 16;
    while (c < (size_t)n) c *= 2;
    if (p.d_tasks) RADIK_HIP(hipFree(p.d_tasks));
    if (p.d_ctl) RADIK_HIP(hipFree(p.d_ctl));
    if (p.d_hist) RADIK_HIP(hipFree(p.d_hist));
    p.d_tasks = nullptr;
    p.d_ctl = nullptr;
    p.d_hist = nullptr;
    p.cap_t = 0;
    if (hipMalloc((void**)&p.d_tasks, c * sizeof(F64Task)) != hipSuccess ||
        hipMalloc((void**)&p.d_ctl, c * sizeof(F64Ctl)) != hipSuccess ||
        hipMalloc((void**)&p.d_hist, c * HIST_BINS * sizeof(uint32_t)) != hipSuccess)
      return set_error(RADIK_ENOMEM, "allocation of the Float64 task state failed");
    RADIK_HIP(hipMemset(p.d_hist, 0, c * HIST_BINS * sizeof(uint32_t)));
    p.cap_t = c;
  }
  if (p.chunks.size() > p.cap_c) {
    RADIK_HIP(hipStreamSynchronize(s));
    size_t c = 64;
    while (c < p.chunks.size()) c *= 2;
    if (p.d_chunks) RADIK_HIP(hipFree(p.d_chunks));
    if (p.d_ccnt) RADIK_HIP(hipFree(p.d_ccnt));
    p.d_chunks = nullptr;
    p.d_ccnt = nullptr;
    p.cap_c = 0;
    if (hipMalloc((void**)&p.d_chunks, c * sizeof(F64Chunk)) != hipSuccess ||
        hipMalloc((void**)&p.d_ccnt, c * sizeof(uint32_t)) != hipSuccess)
      return set_error(RADIK_ENOMEM, "allocation of the Float64 chunk table failed");
    p.cap_c = c;
  }
  // synchronous copies from pageable memory: the vectors may change with the next staging
  if (n) RADIK_HIP(hipMemcpy(p.d_tasks, p.tasks.data(), (size_t)n * sizeof(F64Task), hipMemcpyHostToDevice));
  if (!p.chunks.empty())
    RADIK_HIP(hipMemcpy(p.d_chunks, p.chunks.data(), p.chunks.size() * sizeof(F64Chunk), hipMemcpyHostToDevice));
  return 0;
}

int f64_enqueue(F64Pipe* pp, hipStream_t s) {
  if (!pp || pp->tasks.empty()) return 0;
  F64Pipe& p = *pp;
  const int T = (int)p.tasks.size();
  const unsigned nch = (unsigned)p.chunks.size();
  hipLaunchKernelGGL(k64_init, dim3((T + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, p.d_tasks, p.d_ctl, T);
  RADIK_HIP(hipGetLastError());
  for (int pass = 0; pass < F64_PASSES; ++pass) {
    if (nch) hipLaunchKernelGGL(k64_hist, dim3(nch), dim3(BLOCK), 0, s, p.d_tasks, p.d_chunks, p.d_ctl, p.d_hist, pass);
    hipLaunchKernelGGL(k64_select, dim3(T), dim3(BLOCK), 0, s, p.d_ctl, p.d_hist, pass);
    RADIK_HIP(hipGetLastError());
  }
  if (nch) {
    hipLaunchKernelGGL(k64_ties, dim3(nch), dim3(BLOCK), 0, s, p.d_tasks, p.d_chunks, p.d_ctl, p.d_ccnt);
    hipLaunchKernelGGL(k64_collect, dim3(nch), dim3(BLOCK), 0, s, p.d_tasks, p.d_chunks, p.d_ctl, p.d_ccnt);
    RADIK_HIP(hipGetLastError());
  }
  const size_t lds = (size_t)p.p2max * 12;
  hipLaunchKernelGGL(k64_sort, dim3(T), dim3(BLOCK), lds, s, p.d_tasks, p.d_ctl, p.p2max);
  RADIK_HIP(hipGetLastError());
  return 0;
}

// Opts the sort into LDS for RADIK_F64_MAX_K results (12 bytes each).
int f64_setup(size_t lds_budget) {
  const size_t need = (size_t)RADIK_F64_MAX_K * 12;
  if (need > lds_budget) return 1;
  if (need > 65536 &&
      hipFuncSetAttribute((const void*)k64_sort, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need) != hipSuccess)
    return 1;
  return 0;
}

}  // namespace radik
