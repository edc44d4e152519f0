// radik_sampling.hip -- temperature / top-p (nucleus) sampling from the
// sorted top-k of each task, the step after the top-k path in a serving
// loop (logits rows of a vocabulary).  No counterpart in the reference,
// whose path ends with the sorted [k, T] outputs (src/algorithm.jl:340-367).
//
// Per task, with the results in descending order (largest = true, rev =
// true: v_1 >= v_2 >= ...; equal values by lowest position first):
//   w_i   = exp((v_i - v_1) / T)            (w_i = 1 where v_i == v_1)
//   c_i   = w_1 + ... + w_i                 (float32, one fixed summation order)
//   n     = the first i with c_i >= top_p * c_m   (the nucleus; n = m if none)
//   r     = the first i <= n with c_i > u * c_n  (u: the caller's uniform in [0, 1))
// token = idx_out[r], rank = r - 1 (0-based).  T <= 0 takes r = 1 (greedy).
// The three scans recompute c_i in the same order, so they agree exactly.
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "radik_host.h"

namespace radik {

// Task t's results start at out_off[t]; m[t] of them are valid.
struct SampleArgs {
  const void* vals;
  const void* idx;
  const int64_t* out_off;
  const int32_t* m;
  const float* u;
  int64_t* tokens;
  int32_t* ranks;
  float inv_temp;
  float top_p;
  int key_type;
  int idx_bytes;
  int greedy;
};

__device__ __forceinline__ float value_f32(const void* p, int64_t i, int kt) {
  if (kt == KEY_F16) return __half2float(reinterpret_cast<const __half*>(p)[i]);
  if (kt == KEY_BF16) return __uint_as_float((uint32_t)reinterpret_cast<const uint16_t*>(p)[i] << 16);
  if (kt == KEY_I32) return (float)reinterpret_cast<const int32_t*>(p)[i];
  return reinterpret_cast<const float*>(p)[i];
}

// Inclusive float scan over the block (4 waves of 64): wave shuffles, then the
// wave totals through LDS.  Ends synchronised.
__device__ __forceinline__ float block_scan_f32(float x, float* wsum) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  float pre = 0.0f;
  for (int j = 0; j < w; ++j) pre += wsum[j];
  __syncthreads();
  return x + pre;
}

__global__ __launch_bounds__(BLOCK) void k_topp(SampleArgs a) {
  __shared__ float wsum[BLOCK / 64];
  __shared__ float s_total, s_keep, s_chunk;
  __shared__ int s_first;
  const int t = blockIdx.x;
  const int64_t off = a.out_off[t];
  const int m = a.m[t];
  if (m <= 0) {
    if (threadIdx.x == 0) {
      a.tokens[t] = -1;
      a.ranks[t] = -1;
    }
    return;
  }
  const float v1 = value_f32(a.vals, off, a.key_type);
  auto weight = [&](int i) -> float {
    if (i >= m) return 0.0f;
    const float v = value_f32(a.vals, off + i, a.key_type);
    return v == v1 ? 1.0f : expf((v - v1) * a.inv_temp);
  };
  // the first i (1-based) whose inclusive sum meets pred, scanning in chunks
  // of BLOCK in one fixed order; returns m + 1 if none; sum_at receives c_i
  auto first_where = [&](auto pred, float* sum_at) -> int {
    float carry = 0.0f;
    if (threadIdx.x == 0) s_first = m + 1;
    __syncthreads();
    for (int c0 = 0; c0 < m; c0 += BLOCK) {
      const int i = c0 + (int)threadIdx.x;
      const float c = block_scan_f32(weight(i), wsum) + carry;  // (ends synchronised)
      if (i < m && pred(c)) atomicMin(&s_first, i + 1);
      if (threadIdx.x == BLOCK - 1) s_chunk = c;
      __syncthreads();
      const int f = s_first;
      if (f <= m) {
        if (i + 1 == f && sum_at) *sum_at = c;
        __syncthreads();
        return f;
      }
      carry = s_chunk;
      __syncthreads();  // s_chunk is rewritten by the next chunk
    }
    return m + 1;
  };
  int r = 1;
  if (!a.greedy) {
    // c_m: the total, by the same scan
    float carry = 0.0f;
    for (int c0 = 0; c0 < m; c0 += BLOCK) {
      const float c = block_scan_f32(weight(c0 + (int)threadIdx.x), wsum) + carry;
      if (threadIdx.x == BLOCK - 1) s_chunk = c;
      __syncthreads();
      carry = s_chunk;
      __syncthreads();
    }
    if (threadIdx.x == 0) s_total = carry;
    __syncthreads();
    const float z = s_total;
    const float want = a.top_p * z;
    int n = first_where([&](float c) { return c >= want; }, &s_keep);
    __syncthreads();
    float cn = s_keep;
    if (n > m) {
      n = m;
      cn = z;
    }
    const float th = a.u[t] * cn;
    r = first_where([&](float c) { return c > th; }, nullptr);
    if (r > n) r = n;  // (u * c_n rounds up to c_n)
  }
  if (threadIdx.x == 0) {
    const int64_t j = off + r - 1;
    a.tokens[t] = a.idx_bytes == 8 ? reinterpret_cast<const int64_t*>(a.idx)[j]
                                   : (int64_t)reinterpret_cast<const int32_t*>(a.idx)[j];
    a.ranks[t] = r - 1;
  }
}

int sample_topp(const void* vals, const void* idx, int idx_bytes, const int64_t* out_off, const int32_t* m,
                int64_t num_tasks, int key_type, float temperature, float top_p, const float* u, int64_t* tokens,
                int32_t* ranks, hipStream_t s) {
  if (num_tasks <= 0) return 0;
  SampleArgs a;
  a.vals = vals;
  a.idx = idx;
  a.out_off = out_off;
  a.m = m;
  a.u = u;
  a.tokens = tokens;
  a.ranks = ranks;
  a.greedy = !(temperature > 0.0f);
  a.inv_temp = a.greedy ? 0.0f : 1.0f / temperature;
  a.top_p = top_p;
  a.key_type = key_type;
  a.idx_bytes = idx_bytes;
  hipLaunchKernelGGL(k_topp, dim3((unsigned)num_tasks), dim3(BLOCK), 0, s, a);
  RADIK_HIP(hipGetLastError());
  return 0;
}

}  // namespace radik
