// radik_launch.hip -- the kernel sequence of one key type.  The library
// compiles this file once per key type (RADIK_DT / RADIK_KNS: float32,
// float16, bfloat16, int32); each object holds that type's kernels in its
// own namespace and exports its launch functions to radik_pipeline.hip.
#include <algorithm>

#include "radik_host.h"
#include "radik_kernels.hip"

#define RADIK_PASTE2(a, b) a##b
#define RADIK_PASTE(a, b) RADIK_PASTE2(a, b)

namespace radik {
namespace RADIK_KNS {

// 16-byte loads in flight per lane in the streaming passes: 16-bit keys
// carry twice the elements per load, so their widest setting is 2.
constexpr int UNR_MAX = (EB == 2) ? 2 : 4;

#ifndef RADIK_MED_FULL_CAP
#define RADIK_MED_FULL_CAP 1
#endif

// ---------------------------------------------------------------------------
// Kernel sequence
// ---------------------------------------------------------------------------
// Pass A with the fused selection (fin_cap: see k_finish).
template <int UNR>
static void launch_stream_kernels_hist(const Pipeline& p, uint32_t fin_cap, hipStream_t s) {
  if (p.grid_h <= 0) return;
  if (p.pos_bounds)
    hipLaunchKernelGGL((k_hist<UNR, true, true>), dim3(p.grid_h), dim3(BLOCK), 0, s, p.d_large, p.d_bases,
                       (int)p.n_large, p.d_cstart, p.d_hist, p.d_ctl, p.d_cand, p.d_bmap + p.off_h, fin_cap);
  else
    hipLaunchKernelGGL((k_hist<UNR, true, false>), dim3(p.grid_h), dim3(BLOCK), 0, s, p.d_large, p.d_bases,
                       (int)p.n_large, p.d_cstart, p.d_hist, p.d_ctl, p.d_cand, p.d_bmap + p.off_h, fin_cap);
}

// Pass B over the whole element space (pass A's chunks, one per block);
// blocks skip tasks of at most BIG_LEN elements and those with a usable spill.
template <int UNR>
static void launch_stream_kernels_compact(const Pipeline& p, uint32_t fin_cap, hipStream_t s) {
  if (p.n_chunks <= 0) return;
  hipLaunchKernelGGL(k_compact<UNR>, dim3((unsigned)p.n_chunks), dim3(BLOCK), 0, s, p.d_large, p.d_bases,
                     (int)p.n_large, p.d_cstart, p.d_ctl, p.d_cand, p.d_bmap + p.off_h, BIG_LEN, p.d_shist,
                     (const uint32_t*)p.d_hist, fin_cap);
}

// k_small and k_medium (independent of the long tasks).
static int launch_short_and_medium(const Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s) {
  if (p.n_small) {
    const uint32_t scap = eff_small_cap(f, o);
    hipLaunchKernelGGL(k_small, dim3((unsigned)p.n_small), dim3(BLOCK), (size_t)scap * 16, s, p.d_small, scap);
    RADIK_HIP(hipGetLastError());
  }
  if (p.n_med) {
    // (at least 1/(med_per_cu + 1) of a CU's LDS per block, so a CU keeps
    // room for the sampling and prediction kernels beside the medium tasks)
    const uint32_t sel = std::min<uint32_t>(p.max_m_med, MED_CAP);
    const size_t lds = std::max<size_t>(((size_t)MED_CAP + sel) * 8,
                                        (f.lds_per_block / (o.med_per_cu + 1) + 256) / 8 * 8);
    // every entry of the allocation not reserved for the m results holds
    // winners and candidates (a task with more finishes in global memory)
#if RADIK_MED_FULL_CAP
    const uint32_t cap = (uint32_t)(lds / 8) - sel;
#else
    const uint32_t cap = MED_CAP;
#endif
    const dim3 g((unsigned)p.n_med), b(BLOCK);
    if (p.max_m_med <= (uint32_t)BLOCK) hipLaunchKernelGGL(k_medium<1>, g, b, lds, s, p.d_med, p.d_cand, cap);
    else if (p.max_m_med <= 2u * BLOCK) hipLaunchKernelGGL(k_medium<2>, g, b, lds, s, p.d_med, p.d_cand, cap);
    else if (p.max_m_med <= 4u * BLOCK) hipLaunchKernelGGL(k_medium<4>, g, b, lds, s, p.d_med, p.d_cand, cap);
    else hipLaunchKernelGGL(k_medium<8>, g, b, lds, s, p.d_med, p.d_cand, cap);
    RADIK_HIP(hipGetLastError());
  }
  return 0;
}

static int enqueue_all(Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s, PhaseTimer* tm) {
  const bool timed = tm && tm->on;
  int phase = 0;
  auto mark = [&]() -> int {  // closes the current phase
    if (timed) RADIK_HIP(hipEventRecord(tm->ev[phase], s));
    ++phase;
    return 0;
  };
  int rc;
  if ((rc = mark())) return rc;  // start
  // Short and medium tasks are independent of the long ones: outside the
  // phase timing they run on the side stream, overlapping the sampling and
  // pass A.
  const bool fork = !timed && (p.n_small || p.n_med) && p.n_large && p.side;
  auto fork_side = [&]() -> int {
    RADIK_HIP(hipEventRecord(p.ev_fork, s));
    RADIK_HIP(hipStreamWaitEvent(p.side, p.ev_fork, 0));
    int r = launch_short_and_medium(p, f, o, p.side);
    if (r) return r;
    RADIK_HIP(hipEventRecord(p.ev_join, p.side));
    return 0;
  };
  // fork_late: the side stream starts with the launch (0), once the
  // prediction is done (1: the sampling and prediction have the whole chip
  // and the short and medium tasks overlap pass A only) or once the sampling
  // is done (2, the default: measured best eagerly and as a graph with the
  // tasks staged longest first).
  const int fl = o.fork_late >= 0 ? o.fork_late : 2;
  if (fork && fl == 0 && (rc = fork_side())) return rc;
  const bool join_here =
      o.join_early >= 0 ? o.join_early != 0 : p.total >= p.total_med + p.total_small;  // (before k_finish)
  if (p.n_large) {
    // k_finish's LDS: cap entries for a task's winners + candidates (a power
    // of two >= 2048, which also holds the 4096-bin histogram of the wide
    // digits), plus the m results of fin tasks; results sorted in k_finish
    // when every task's m fits cap (else k_sort follows, without fin tasks).
    const uint32_t budget = (uint32_t)((f.lds_per_block - FINISH_STATIC_LDS) / 8);
    uint32_t cap = 2048;
    {
      const uint32_t want = (p.n_large <= (int64_t)f.cus) ? 16384u : std::max<uint32_t>(2048u, eff_refine_cap(f, o));
      while (cap * 2 <= want && cap * 2 <= budget) cap *= 2;
    }
    bool sort_here = false;
    uint32_t sel = 0;
    for (;;) {
      sort_here = p.max_m <= cap && p.max_m <= (uint32_t)BLOCK * 64;
      sel = sort_here ? std::min<uint32_t>(p.max_m, cap) : 0u;
      if (cap + sel <= budget || cap <= 2048) break;
      cap /= 2;
    }
    if (cap + sel > budget) {
      sort_here = false;
      sel = 0;
    }
    const uint32_t fin_cap = sort_here ? cap : 0u;
    uint32_t emax = 1;
    while (sort_here && (uint32_t)BLOCK * emax < p.max_m) emax *= 2;
    {
      if (p.total_groups > 0) {
        hipLaunchKernelGGL(k_sample, dim3(p.grid_s), dim3(BLOCK), 0, s, p.d_large, p.d_sbases, (int)p.n_large,
                           p.total_groups, p.d_shist, p.d_bmap + p.off_s);
        RADIK_HIP(hipGetLastError());
      }
      if ((rc = mark())) return rc;
      if (fork && fl == 2 && (rc = fork_side())) return rc;
      hipLaunchKernelGGL(k_predict, dim3((unsigned)p.n_large), dim3(BLOCK), 0, s, p.d_large, p.d_sbases,
                         p.d_shist, p.d_ctl);
      RADIK_HIP(hipGetLastError());
      if ((rc = mark())) return rc;
      if (fork && fl == 1 && (rc = fork_side())) return rc;
      switch (o.unr) {
        case 1: launch_stream_kernels_hist<1>(p, fin_cap, s); break;
        case 4: launch_stream_kernels_hist<UNR_MAX>(p, fin_cap, s); break;
        default: launch_stream_kernels_hist<2>(p, fin_cap, s); break;
      }
    }
    RADIK_HIP(hipGetLastError());
    if ((rc = mark())) return rc;
    // Tasks spread over several pass-A blocks are selected by k_finish, and
    // those longer than BIG_LEN by k_compact's blocks (pass B over several
    // blocks re-reads the ones without a usable spill; k_finish re-reads the
    // shorter ones).
    if ((rc = mark())) return rc;
    if (p.has_big) {
      switch (o.unr) {
        case 1: launch_stream_kernels_compact<1>(p, fin_cap, s); break;
        case 4: launch_stream_kernels_compact<UNR_MAX>(p, fin_cap, s); break;
        default: launch_stream_kernels_compact<2>(p, fin_cap, s); break;
      }
      RADIK_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_split, dim3((unsigned)p.n_split), dim3(BLOCK), 0, s, p.d_large, p.d_ctl, p.d_cand,
                         p.d_xmap, p.d_shist);
      RADIK_HIP(hipGetLastError());
    }
    if ((rc = mark())) return rc;
    // The side stream is waited for here, not at the end, when its tasks are
    // expected to end before pass A: the wait is then resolved while pass A
    // runs, and the next launch starts right behind k_finish.
    if (fork && join_here) RADIK_HIP(hipStreamWaitEvent(s, p.ev_join, 0));
    {
      const size_t lds = ((size_t)cap + sel) * 8;
      const dim3 b(BLOCK);
      const int sh = sort_here ? 1 : 0;
      const int sp = p.has_big ? 1 : 0;  // k_split listed the long tasks
      // (extra blocks sort the rank tail's runs; k_finish's LDS holds a run)
      const dim3 g((unsigned)(p.n_large + p.n_rt));
      const int nf = (int)p.n_large;
      switch (emax) {
        case 1: hipLaunchKernelGGL(k_finish<1>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        case 2: hipLaunchKernelGGL(k_finish<2>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        case 4: hipLaunchKernelGGL(k_finish<4>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        case 8: hipLaunchKernelGGL(k_finish<8>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        case 16: hipLaunchKernelGGL(k_finish<16>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        case 32: hipLaunchKernelGGL(k_finish<32>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
        default: hipLaunchKernelGGL(k_finish<64>, g, b, lds, s, p.d_large, p.d_ctl, p.d_hist, p.d_cand, cap, fin_cap, BIG_LEN, sh, sp, nf, p.d_rmap, p.d_shist); break;
      }
      RADIK_HIP(hipGetLastError());
    }
    if ((rc = mark())) return rc;
    // the rank tail's ranks (after its run sorts in k_finish's launch):
    // beside k_sort_chunks in one launch when that sort runs
    const uint32_t fit = (uint32_t)(f.lds_per_block / 8 / RT_RUN) * RT_RUN;
    const uint32_t le =
        std::max<uint32_t>(RT_RUN, std::min<uint32_t>(std::min<uint32_t>(p.rt_lds, RT_RANK_LDS), fit));
    const unsigned nrank = (unsigned)(p.n_rt * (RT_RUN / RT_ENTRIES));
    const bool chunk_sort = !sort_here && p.max_m > 4096 && p.max_m <= SORT_CHUNK * SORT_MAX_RUNS;
    if (p.n_rt > 0 && !chunk_sort) {
      hipLaunchKernelGGL(k_rt_rank, dim3(nrank), dim3(BLOCK), (size_t)le * 8, s, p.d_large, p.d_ctl,
                         (const uint64_t*)p.d_cand, p.d_rmap, le);
      RADIK_HIP(hipGetLastError());
    }
    if (!sort_here) {
      if (chunk_sort) {
        // several blocks per task: sorted runs, then ranks across runs
        if (p.n_rt > 0)
          hipLaunchKernelGGL(k_rank_sort, dim3(nrank + (unsigned)p.total_runs), dim3(BLOCK),
                             (size_t)std::max<uint32_t>(le, SORT_CHUNK) * 8, s, p.d_large, p.d_ctl, p.d_cand,
                             p.d_rmap, le, nrank, p.d_cbases, (int)p.n_large);
        else
          hipLaunchKernelGGL(k_sort_chunks, dim3((unsigned)p.total_runs), dim3(BLOCK), 0, s, p.d_large, p.d_ctl,
                             p.d_cbases, (int)p.n_large, p.d_cand);
        RADIK_HIP(hipGetLastError());
        const unsigned rg = (unsigned)((p.total_m + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL(k_sort_rank, dim3(rg), dim3(BLOCK), 0, s, p.d_large, p.d_ctl, p.d_mbases,
                           (int)p.n_large, p.total_m, (const uint64_t*)p.d_cand);
      } else {
        uint32_t e = 1;
        while ((uint32_t)BLOCK * e < p.max_m && (uint32_t)BLOCK * e < f.sort_cap) e *= 2;
        const size_t lds = (size_t)BLOCK * e * 8;
        const dim3 g((unsigned)p.n_large), b(BLOCK);
        switch (e) {
          case 1: hipLaunchKernelGGL(k_sort<1>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          case 2: hipLaunchKernelGGL(k_sort<2>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          case 4: hipLaunchKernelGGL(k_sort<4>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          case 8: hipLaunchKernelGGL(k_sort<8>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          case 16: hipLaunchKernelGGL(k_sort<16>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          case 32: hipLaunchKernelGGL(k_sort<32>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
          default: hipLaunchKernelGGL(k_sort<64>, g, b, lds, s, p.d_large, p.d_ctl, p.d_cand); break;
        }
      }
      RADIK_HIP(hipGetLastError());
    }
  } else {
    for (int i = 0; i < 6; ++i)
      if ((rc = mark())) return rc;
  }
  if ((rc = mark())) return rc;  // end of sort
  if (fork) {
    if (!join_here) RADIK_HIP(hipStreamWaitEvent(s, p.ev_join, 0));
  } else if ((rc = launch_short_and_medium(p, f, o, s))) {
    return rc;
  }
  if ((rc = mark())) return rc;  // end of small
  return 0;
}

}  // namespace RADIK_KNS

// The whole kernel sequence of a launch of this key type (see enqueue_all).
int RADIK_PASTE(enqueue_, RADIK_KNS)(Pipeline& p, const DevFacts& f, const LaunchOpts& o, hipStream_t s,
                                    PhaseTimer* tm) {
  return RADIK_KNS::enqueue_all(p, f, o, s, tm);
}

// Opts this key type's per-task kernels into lds_budget bytes of LDS per
// block (when above 64 KiB) and reports pass A's resident blocks per CU.
int RADIK_PASTE(launch_setup_, RADIK_KNS)(size_t lds_budget, int* hist_occ) {
  using namespace RADIK_KNS;
  if (lds_budget > 65536) {
    const int want = (int)lds_budget;
    const int fin_want = want - (int)FINISH_STATIC_LDS;
    const hipFuncAttribute A = hipFuncAttributeMaxDynamicSharedMemorySize;
    bool ok = hipFuncSetAttribute((const void*)k_finish<1>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<2>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<4>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<8>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<16>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<32>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_finish<64>, A, fin_want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_rt_rank, A, want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_rank_sort, A, want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_sort<32>, A, want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_sort<64>, A, want) == hipSuccess &&
              hipFuncSetAttribute((const void*)k_small, A, want - 20 * 1024) == hipSuccess;
    if (!ok) return 1;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_hist<UNR_MAX, true, true>, BLOCK, 0) !=
          hipSuccess ||
      occ < 1) {
    (void)hipGetLastError();
    occ = 4;
  }
  *hist_occ = occ;
  return 0;
}

#if RADIK_DT == 0
// Diagnostics (float32 only): the pass-level intermediates the parity tests
// inspect, the scans, and the copy kernel that calibrates traffic counters.
using namespace RADIK_KNS;

// Diagnostics entry points' kernels.
int diag_scan(const uint32_t* in, uint32_t* ow, uint32_t* ob, int64_t n, hipStream_t s) {
  if (n <= 0) return 0;
  const int grid = (int)((n + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_diag_scan, dim3(grid), dim3(BLOCK), 0, s, in, ow, ob, n);
  RADIK_HIP(hipGetLastError());
  return 0;
}

int diag_copy(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  const int64_t n16 = bytes / 16;
  if (n16 <= 0) return 0;
  int dev = 0, cus = 0;
  RADIK_HIP(hipGetDevice(&dev));
  RADIK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int64_t want = (int64_t)(cus > 0 ? cus : 256) * 16;
  const int grid = (int)std::min<int64_t>(want, (n16 + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_diag_copy, dim3(grid), dim3(BLOCK), 0, s, (const f32x4*)src, (f32x4*)dst, n16);
  RADIK_HIP(hipGetLastError());
  return 0;
}


int diag_pass1(Pipeline& p, const DevFacts& f, uint32_t* hist_out, uint32_t* sel_out, hipStream_t s) {
  if (!p.n_large) return 0;
  // Control blocks without a prediction: pass A only counts.
  // staged without prediction: k_predict only initialises the control blocks
  hipLaunchKernelGGL(k_predict, dim3((unsigned)p.n_large), dim3(BLOCK), 0, s, p.d_large, p.d_sbases, p.d_shist,
                     p.d_ctl);
  RADIK_HIP(hipGetLastError());
  if (p.grid_h > 0) {
    hipLaunchKernelGGL((k_hist<2, false, false>), dim3(p.grid_h), dim3(BLOCK), 0, s, p.d_large, p.d_bases,
                       (int)p.n_large, p.d_cstart, p.d_hist, p.d_ctl, p.d_cand, p.d_bmap + p.off_h, 0u);
    RADIK_HIP(hipGetLastError());
  }
  RADIK_HIP(hipMemcpyAsync(hist_out, p.d_hist, (size_t)p.n_large * HIST_BINS * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_select, dim3((unsigned)p.n_large), dim3(BLOCK), 0, s, p.d_large, p.d_hist, p.d_ctl, 0u);
  RADIK_HIP(hipGetLastError());
  RADIK_HIP(hipMemcpy2DAsync(sel_out, 4 * sizeof(uint32_t), p.d_ctl, sizeof(TaskCtl), 4 * sizeof(uint32_t),
                             (size_t)p.n_large, hipMemcpyDeviceToDevice, s));
  return 0;
}

#endif

}  // namespace radik
