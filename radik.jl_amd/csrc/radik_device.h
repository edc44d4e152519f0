// radik_device.h -- layouts and device helpers shared by the gfx950 kernels
// and the host orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace radik {

// Per-task flag bits carried in TaskDesc::flags (the launch-level C-ABI flags
// plus the index width).
constexpr uint32_t F_LARGEST = 0x01u;
constexpr uint32_t F_ASCEND = 0x02u;
constexpr uint32_t F_WITHSCALE = 0x04u;
constexpr uint32_t F_WITHIDXIN = 0x08u;
constexpr uint32_t F_IDX64 = 0x100u;

// Key types.  The kernels are compiled once per key type (RADIK_DT selects
// it in a kernel translation unit); a launch holds tasks of one type.
constexpr int KEY_F32 = 0;   // float32 (the reference's only type)
constexpr int KEY_F16 = 1;   // IEEE binary16
constexpr int KEY_BF16 = 2;  // bfloat16
constexpr int KEY_I32 = 3;   // int32
constexpr int NUM_KEY_TYPES = 4;  // key types of the 32-bit-key path
// Float64 keys take a path of their own (radik_f64.hip): their key needs all
// 64 bits of a composite.
constexpr int KEY_F64 = 4;
inline int key_bytes(int kt) { return (kt == KEY_F16 || kt == KEY_BF16) ? 2 : (kt == KEY_F64 ? 8 : 4); }

constexpr uint32_t NO_BIN = 0xFFFFFFFFu;
constexpr int HIST_BINS = 4096;  // first digit: rank-key bits [31:20]
constexpr int BLOCK = 256;       // threads of every kernel in this library
// Predictions are refined to a composite bound only for samples of at most
// this many groups (the host picks pass A's variant by it).
constexpr int PRED_MAX_G = 64;
// Sampling: groups of SAMPLE_GROUP consecutive elements, SAMPLE_GROUPS_PER_BLOCK
// per block of k_sample.
constexpr int SAMPLE_GROUP = BLOCK;
constexpr int SAMPLE_GROUPS_PER_BLOCK = 16;
// Multi-block sort of large result sets: runs of SORT_CHUNK, at most
// SORT_MAX_RUNS of them per task.
constexpr uint32_t SORT_CHUNK = BLOCK * 8;
constexpr uint32_t SORT_MAX_RUNS = 64;
// Rank tail of tasks over 2^20 elements with a covering spill: their listed
// winners and candidates are sorted in runs of RT_RUN (one block each), then
// each entry's rank over all runs (binary searches in LDS) gives both the
// selection (rank < m) and its output slot.  At most RT_MAX_RUNS runs.
constexpr uint32_t RT_RUN = BLOCK * 2;
constexpr int64_t RT_MAX_RUNS = 256;
// Entries of the runs a rank block holds in LDS at once (the launch's
// longest list up to RT_LDS, and at most RT_RANK_LDS).
constexpr uint32_t RT_LDS = 16384;
constexpr uint32_t RT_RANK_LDS = 8192;

// One task of a launch.  Pointers address the task's own first element, so a
// launch may mix tasks from different buffers (the serving engine does).
struct TaskDesc {
  const void* in;       // len values of the launch's key type
  const void* idx_in;   // len indices (int32/int64) when F_WITHIDXIN
  void* val_out;        // k values of the key type
  void* idx_out;        // k indices (int32/int64)
  int64_t len;          // elements (< 2^32)
  int64_t base;         // start in the launch's concatenated element space;
                        // also the task's region of the candidate buffer
  int32_t k;            // requested k (>= 0)
  uint32_t flags;       // F_* bits
  int32_t nblk;         // pass-A chunks overlapping the task
  int32_t rt_runs;      // rank-tail runs the launch gives the task (0: none)
};
static_assert(sizeof(TaskDesc) == 64, "TaskDesc layout");

// Per-task state between the passes: k_predict initialises it, the
// first-digit selection (end of pass A) fills it.
struct TaskCtl {
  uint32_t b_win;     // first-digit values < b_win are winners
  uint32_t b_cand;    // first-digit value of the candidates, NO_BIN if none
  uint32_t krem;      // candidates still to take (0: none)
  uint32_t ncand;     // candidates expected (= count of bin b_cand)
  uint32_t nwin;      // winners expected (= m - krem)
  uint32_t m;         // min(k, len): valid outputs
  uint32_t win_ctr;   // slot counter for winners (compaction)
  uint32_t cand_ctr;  // entries in the task's spill region
  uint32_t b_pred;    // predicted first digit (NO_BIN: no prediction, no spill)
  uint32_t spill_all; // 1: the spill holds the m best; 0: the task is re-read
  uint32_t fin;       // 1: the m best are gathered in k_finish's LDS (their source fits it)
  uint32_t done;      // 1: the first digit is selected
  uint32_t list_ctr;  // candidates listed by k_split, or by pass B over several blocks
  uint32_t r_pred;    // pass A spills the composites <= (r_pred << 32 | p_pred)
  uint32_t p_pred;
  uint32_t rt;        // 1: the rank tail (run sorts in k_finish's launch, k_rt_rank) finishes the task
};
static_assert(sizeof(TaskCtl) == 64, "TaskCtl layout");

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)

// The task pointers (TaskDesc) address device global memory, but as fields
// loaded from memory they reach the compiler as generic pointers, whose
// (flat) accesses count against the LDS counter too and make every LDS wait
// drain the outstanding loads.  Accesses through them go through glb().
#define RADIK_AS1 __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ RADIK_AS1 T* glb(T* p) {
  return (RADIK_AS1 T*)p;
}

// Order-preserving key of a float: unsigned order of key == numeric order,
// with -0 < +0 and NaNs outside +-Inf (RadiK.jl src/utils.jl:29-33).
__device__ __forceinline__ uint32_t fkey(float x) {
  uint32_t u = __float_as_uint(x);
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}

__device__ __forceinline__ float fkey_inv(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
  return __uint_as_float(u);
}

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Lanes of mask below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return __builtin_amdgcn_ballot_w64(p);
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

// Inclusive prefix sum across the 64 lanes of a wave with DPP row shifts and
// the GFX9 row broadcasts: rows of 16 first, then row 0 -> 1, 2 -> 3, and
// lane 31 -> rows 2 and 3.  Lanes of disabled rows or without a source read 0.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

// Inclusive scan over a block of BLOCK threads (4 waves): wave scan, then the
// wave totals through LDS.  wsum: LDS scratch of BLOCK/64 words.  Ends with
// the block synchronised.
__device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t* wsum) {
  const uint32_t tid = threadIdx.x;
  uint32_t incl = wave_incl_scan(v);
  if ((tid & 63u) == 63u) wsum[tid >> 6] = incl;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t w = 0; w < (tid >> 6); ++w) pre += wsum[w];
  __syncthreads();
  return incl + pre;
}

// Finds the bin b with before(b) < need <= before(b) + count(b), where each
// thread holds PER consecutive bins (thread i: bins i*PER .. i*PER+PER-1) and
// before(b) sums all lower bins.  Requires 1 <= need <= total.  res (LDS, 3
// words) receives b, before(b), count(b); the block is synchronised on return.
template <int PER>
__device__ __forceinline__ void block_find_bin(const uint32_t (&c)[PER], uint32_t need,
                                               uint32_t* wsum, uint32_t* res) {
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) s += c[j];
  const uint32_t incl = block_incl_scan(s, wsum);
  const uint32_t excl = incl - s;
  if (excl < need && need <= incl) {
    uint32_t before = excl;
    bool done = false;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      if (!done && before + c[j] >= need) {
        res[0] = threadIdx.x * PER + j;
        res[1] = before;
        res[2] = c[j];
        done = true;
      }
      before += c[j];
    }
  }
  __syncthreads();
}

#endif  // device

}  // namespace radik
