/*
 * radik.h -- C-ABI of libradik_hip.so, the MI355X (gfx950) radix top-k library.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device addresses;
 * hipStream_t values are passed as void* (NULL = the default stream).
 * Every function returns 0 on success or one of the RADIK_E* codes below, and
 * leaves a message for radik_last_error() (per calling thread).
 *
 * The Julia shim (radik.jl_amd/julia/RadiK/src/RadiK.jl) binds these with
 * ccall; the Python mirror (radik.jl_amd/radik) with ctypes.  Each entry cites
 * the reference interface (WilliBee/RadiK.jl) it replaces.
 */
#ifndef RADIK_H
#define RADIK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flag bits: the Val switches of topk_radix_select! (src/algorithm.jl:170-175). */
#define RADIK_LARGEST     0x01u /* Val{LARGEST}: top-k by largest value             */
#define RADIK_ASCEND      0x02u /* Val{ASCEND}: output sorted ascending (= !rev)    */
#define RADIK_WITHSCALE   0x04u /* Val{WITHSCALE}: select on x - first element      */
#define RADIK_WITHIDXIN   0x08u /* Val{WITHIDXIN}: idx_out = idx_in[pos]            */
#define RADIK_WITHPACKING 0x10u /* Val{WITHPACKING}: vector loads (vload_bytes)    */

/* Key type, bits 8-9 of the same flags word (no counterpart in the
 * reference, which takes Float32 only: README.md:18-24,
 * algorithm.jl:492-494).  The element arrays (val_in, val_out) of a call
 * then hold that type; pass them cast to the declared pointer type.
 * Ordering: the float types by their IEEE-754 total order (-NaN < -Inf <
 * ... < -0 < +0 < ... < +Inf < +NaN), Int32 by value.  WITHSCALE takes
 * Float32 keys only.  Float64 keys (bits 8-10 = 4) run through the
 * workspace path (radik_topk_radix_select) and radik_topk_typed only, with
 * k <= RADIK_F64_MAX_K; plans and the serving engine reject them. */
#define RADIK_KEY_SHIFT 8
#define RADIK_KEY_F32  0x000u /* Float32 (default)              */
#define RADIK_KEY_F16  0x100u /* Float16 (IEEE binary16)         */
#define RADIK_KEY_BF16 0x200u /* BFloat16                        */
#define RADIK_KEY_I32  0x300u /* Int32                           */
#define RADIK_KEY_F64  0x400u /* Float64                         */
#define RADIK_KEY_MASK 0x700u
#define RADIK_F64_MAX_K 8192

/* Status codes. */
#define RADIK_OK            0
#define RADIK_EINVAL        1 /* bad argument (k < 0, null pointer, bad flags)       */
#define RADIK_ESIZE         2 /* output size mismatch (algorithm.jl:498-501)        */
#define RADIK_EINDICES      3 /* indices length != data length (algorithm.jl:507-509) */
#define RADIK_EWORKSPACE    4 /* workspace too small: max(task_lens) > n or T > num_tasks */
#define RADIK_EHIP          5 /* HIP runtime error                                   */
#define RADIK_ENOMEM        6 /* device or pinned allocation failed                  */
#define RADIK_EADMIT        7 /* serving: request exceeds the workspace budget       */
#define RADIK_ENOTREADY     8 /* serving: result not complete yet                    */
#define RADIK_ELIMIT        9 /* radik_topk_noalloc: input beyond its fixed storage  */

typedef struct radik_ws radik_ws;
typedef struct radik_plan radik_plan;
typedef struct radik_server radik_server;

/* Library identity and errors. */
const char *radik_version(void);
const char *radik_last_error(void);

/* RadiKWorkspace(backend, n, num_tasks, ::Type{I}=Int32)   src/algorithm.jl:58-83.
 * n: longest task the workspace serves; num_tasks: most tasks per call;
 * counter_bytes: 4 (Int32) or 8 (Int64), sizeof(I).  Device buffers are
 * allocated on the current device; the spill buffer grows on first use. */
int radik_ws_create(radik_ws **ws, int64_t n, int64_t num_tasks, int counter_bytes, void *stream);
int radik_ws_destroy(radik_ws *ws);

/* topk_radix_select!(val_out, idx_out, ws, val_in, idx_in, task_lens, k,
 *     Val(LARGEST), Val(ASCEND), Val(WITHSCALE), Val(WITHIDXIN),
 *     Val(WITHPACKING), Val(VLOAD_SIZE_BYTES))      src/algorithm.jl:162-368.
 * val_in: concatenated tasks (device, sum(task_lens) floats); idx_in: same
 * length, int32 or int64 (idx_bytes) when RADIK_WITHIDXIN, else ignored;
 * task_lens: HOST array of num_tasks lengths; val_out/idx_out: device [k,
 * num_tasks] column-major.  Outputs per task: min(k, len) sorted results,
 * then 0.0f / -1 padding.  Synchronises the stream before returning, like
 * the reference (algorithm.jl:353).  Each task holds fewer than 2^31
 * elements with Int32 indices and fewer than 2^32 with Int64 (positions are
 * 32-bit inside the selection keys; the reference's Int64 counters,
 * algorithm.jl:150-154, reach further): longer tasks give RADIK_EINVAL. */
int radik_topk_radix_select(radik_ws *ws, float *val_out, void *idx_out, int idx_bytes,
                            const float *val_in, const void *idx_in, const int64_t *task_lens,
                            int64_t num_tasks, int64_t k, uint32_t flags, int vload_bytes, void *stream);

/* topk!(val_out, idx_out, data, task_lens, k; indices, largest, rev)
 * src/algorithm.jl:477-522: validates output sizes (out_rows x out_cols must
 * be k x num_tasks) and the indices length (n_indices must be 0 or
 * sum(task_lens)), allocates a workspace for this call (as the reference
 * does, algorithm.jl:503), runs with WITHPACKING and 32-byte loads. */
int radik_topk(float *val_out, void *idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
               const float *val_in, const void *indices, int64_t n_indices, const int64_t *task_lens,
               int64_t num_tasks, int64_t k, int largest, int rev, void *stream);

/* topk! for any key type: key_type 0 Float32, 1 Float16, 2 BFloat16,
 * 3 Int32, 4 Float64 (RADIK_KEY_* >> RADIK_KEY_SHIFT); val_in / val_out
 * hold that type.  Otherwise radik_topk (src/algorithm.jl:477-522
 * generalised past the reference's Float32-only check,
 * algorithm.jl:492-494). */
int radik_topk_typed(int key_type, void *val_out, void *idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
                     const void *val_in, const void *indices, int64_t n_indices, const int64_t *task_lens,
                     int64_t num_tasks, int64_t k, int largest, int rev, void *stream);

/* Allocation-free variant of radik_topk.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, hosts that forbid hipMalloc).  Same
 * arguments, validation and results as radik_topk, but the device work
 * space is a fixed array inside the library (one copy per GPU, 50 MiB, held
 * until the process ends) and the call allocates nothing on the device or
 * the host.  Limits (RADIK_ELIMIT beyond them): num_tasks <=
 * RADIK_NOALLOC_MAX_TASKS, sum(task_lens) <= RADIK_NOALLOC_MAX_ELEMS, at
 * most RADIK_NOALLOC_MAX_LONG tasks longer than 2048 elements.  Calls are
 * serialised inside the library and synchronise the stream before
 * returning; every call on a GPU reuses that GPU's array (the library
 * clears what must start at zero), so work on it must not overlap. */
#define RADIK_NOALLOC_MAX_TASKS 8192
#define RADIK_NOALLOC_MAX_ELEMS 4194304
#define RADIK_NOALLOC_MAX_LONG 512
int radik_topk_noalloc(float *val_out, void *idx_out, int idx_bytes, int64_t out_rows, int64_t out_cols,
                       const float *val_in, const void *indices, int64_t n_indices, const int64_t *task_lens,
                       int64_t num_tasks, int64_t k, int largest, int rev, void *stream);

/* Persistent batched form (no reference counterpart: the reference launches
 * one grid per pass, algorithm.jl:221-334).  A plan fixes task lengths, a k
 * per task and flags; radik_plan_execute enqueues the whole pass sequence
 * asynchronously (eager launches, or the replay of a captured hipGraph
 * after radik_plan_set_graph(plan, 1)).  Task t reads val_in at the
 * sum of earlier lengths and writes k_per_task[t] outputs to val_out/idx_out
 * at the sum of earlier k's (equal k gives the [k, T] layout). */
int radik_plan_create(radik_plan **plan, const int64_t *task_lens, const int64_t *k_per_task,
                      int64_t num_tasks, uint32_t flags, int idx_bytes, void *stream);
int radik_plan_execute(radik_plan *plan, void *val_out, void *idx_out, const void *val_in,
                       const void *idx_in, void *stream);
/* use_graph = 1 captures the sequence into a hipGraph and replays it, 0 (the
 * default: faster on MI355X for every measured workload) launches the
 * kernels eagerly; same results. */
int radik_plan_set_graph(radik_plan *plan, int use_graph);
/* Tuning knobs (results do not depend on them): vload_bytes 16/32/64 (bytes
 * in flight per lane per step of the streaming passes, 0 = keep), streaming
 * blocks per CU (0 = keep; up to the pass's resident blocks per CU: equal
 * chunks, one per block; above it: chunks shrinking towards the end of the
 * pass), longest task run whole in one block, candidates
 * held in LDS by the finishing kernel of launches with many tasks (0 = device
 * defaults), and whether long tasks are sampled so one read of the input
 * suffices. */
int radik_plan_set_options(radik_plan *plan, int vload_bytes, int blocks_per_cu, int small_cap, int refine_cap,
                           int predict);
/* One tuning knob by name (results do not depend on them): "vload_bytes"
 * (16/32/64), "blocks_per_cu", "small_cap", "refine_cap", "predict" (0/1)
 * as in radik_plan_set_options, and "medium" (0/1: tasks of 2^11..2^16
 * elements run whole in one block beside the streaming passes, or join
 * the streaming passes and are re-read by the finishing kernel), and
 * "fork_late" (0/1/2: the short and medium tasks' kernels start with the
 * launch, once the long tasks' prediction is done, or once their sampling is
 * done; -1, the default: 2), "join_early"
 * (1: the plan's stream waits for the short and medium tasks before the
 * finishing kernel, 0: at the end of the launch; -1, the default: 1 when the
 * long tasks hold at least as many elements as the others), "longest_first"
 * (1, the default: tasks are staged longest first, so the one-block-per-task
 * kernels start their slowest blocks first; 0: in the caller's order),
 * "sample_min_len" (shortest task that is
 * sampled and gets a spill bound, default 32768; shorter ones run in
 * k_medium) and
 * "medium_per_cu"
 * (1..16: how many medium-task blocks may share a CU; each reserves that
 * share of the CU's LDS). */
int radik_plan_set_option(radik_plan *plan, const char *name, int64_t value);
/* Per-kernel timing: when enabled, execute runs eagerly (every kernel on the
 * plan's stream) and records HIP events between the phases; radik_plan_timing
 * returns the accumulated milliseconds of the 8 phases (sample, predict,
 * hist = pass A with the selection of tasks inside one of its blocks, select
 * = no kernel of its own (the phase keeps its slot; only the gap between two
 * event records), compact = the selection, re-read and listing of tasks over
 * 2^20 elements when present, finish, sort = only for results beyond the
 * finishing kernel's LDS and the rank tail's ranks, small = short and medium
 * tasks) and the number of executes timed. */
int radik_plan_set_timing(radik_plan *plan, int enable);
int radik_plan_timing(radik_plan *plan, double *ms8, int64_t *count);
/* Temperature / top-p (nucleus) sampling from the sorted top-k of each task
 * (no counterpart in the reference, whose path ends with the sorted [k, T]
 * outputs, src/algorithm.jl:340-367): the step after the top-k in a serving
 * loop over vocabulary logits.  Enqueued on `stream` after the plan's
 * execute (no host synchronisation in between); reads the outputs that
 * execute wrote (val_out / idx_out, descending: the plan must have
 * largest = true and rev = true).  Per task, over its m = min(k, len)
 * results v_1 >= v_2 >= ... (float32 arithmetic, one fixed summation order):
 * w_i = exp((v_i - v_1) / temperature), c_i = w_1 + ... + w_i, the nucleus
 * n = the first i with c_i >= top_p * c_m (m if none), and the sample r =
 * the first i <= n with c_i > uniforms[t] * c_n.  tokens[t] = idx_out[r]
 * (int64), ranks[t] = r - 1 (0-based); temperature <= 0 takes r = 1
 * (greedy; uniforms unused); empty tasks give -1.  Device arrays:
 * uniforms (num_tasks floats in [0, 1)), tokens, ranks.  Logits must not be
 * NaN. */
int radik_plan_sample(radik_plan *plan, const void *val_out, const void *idx_out, float temperature, float top_p,
                      const float *uniforms, int64_t *tokens, int32_t *ranks, void *stream);

/* How the last execute staged the plan's tasks (after the first execute):
 * info[0..8] = long tasks (the streaming passes), their elements (what
 * pass A reads), medium tasks, their elements, short tasks, their elements,
 * pass-A chunks (its grid), sample groups, whether some task exceeds 2^20
 * elements.  n: entries wanted (up to RADIK_PLAN_INFO_LEN). */
#define RADIK_PLAN_INFO_LEN 9
int radik_plan_info(radik_plan *plan, int64_t *info, int n);
int radik_plan_destroy(radik_plan *plan);

/* Serving engine (no reference counterpart): continuous batching of top-k
 * requests against a device workspace budget.  Requests are row-major
 * matrices on the device (rows x cols floats); each row is one task.
 * submit never blocks on the GPU; poll reports completion; results land in
 * the caller's val_out/idx_out (rows x k, row-major, int32 indices,
 * 1-based column positions). */
int radik_server_create(radik_server **srv, int64_t budget_bytes, int max_inflight, uint32_t flags);
int radik_server_submit(radik_server *srv, const void *in, int64_t rows, int64_t cols, int64_t k,
                        void *val_out, int32_t *idx_out, int64_t *ticket);
/* Returns RADIK_OK when the ticket's results are complete (latency_ms: host
 * submit -> completion event), RADIK_ENOTREADY otherwise. */
int radik_server_poll(radik_server *srv, int64_t ticket, double *latency_ms);
int radik_server_wait(radik_server *srv, int64_t ticket, double *latency_ms);
int radik_server_drain(radik_server *srv);
int radik_server_stats(radik_server *srv, int64_t *launches, int64_t *rows_done, int64_t *max_coalesced);
int radik_server_destroy(radik_server *srv);

/* Pass-level diagnostics for parity tests (no reference counterpart; they
 * expose what the reference's kernel unit tests inspect, test/test_count_bin.jl
 * and test/test_select_bin.jl).  Runs pass 1 (histogram of the rank-key digit
 * [31:20]) and the bin selection for the tasks; hist_out receives num_tasks x
 * 4096 uint32 counts (device), sel_out num_tasks x 4 uint32 (device): winner
 * bound, candidate bin (0xFFFFFFFF: none), residual k, candidate count. */
int radik_diag_pass1(const float *val_in, const int64_t *task_lens, int64_t num_tasks, int64_t k,
                     uint32_t flags, uint32_t *hist_out, uint32_t *sel_out, void *stream);
/* Wave64 inclusive scan (DPP) of n uint32 values per 64-lane wave and the
 * 256-thread block scan built on it, as the selection kernels use them
 * (counterpart of @warpreduce, src/utils.jl:84-100): out_wave[i] = inclusive
 * scan within i's wave, out_block[i] = inclusive scan within i's block of 256. */
int radik_diag_scan(const uint32_t *in, uint32_t *out_wave, uint32_t *out_block, int64_t n, void *stream);
/* Host logic without a device call (for CPU tests): pass A's chunk table
 * over `total` elements (nchunks chunks, starts[0..nchunks] when cap is
 * enough), the routing of tasks as a launch stages them (cls: 0 short,
 * 1 medium, 2 long with a sample of groups[t] groups, 3 long without), the
 * serving engine's next batch (how many of the pending footprints it takes
 * given the slot share, the pool in use and the budget) and a request's
 * footprint. */
int radik_diag_plan_chunks(int64_t total, int64_t blocks, int64_t spread, int guided, int64_t gran,
                           int64_t min_chunk, int64_t *starts, int64_t cap, int64_t *nchunks);
int radik_diag_route(const int64_t *task_lens, const int64_t *k_per_task, int64_t num_tasks, int small_cap,
                     int predict, int medium, int32_t *cls, int64_t *groups);
int radik_diag_batch(const int64_t *bytes, int64_t n, int64_t share, int64_t pool_used, int64_t budget,
                     int64_t *take);
int radik_diag_footprint(int64_t rows, int64_t cols, int64_t *bytes);
/* A plain 16-byte streaming copy of `bytes` (multiple of 16): the known-size
 * reference that scales the profiler's HBM traffic counters. */
int radik_diag_copy(void *dst, const void *src, int64_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* RADIK_H */
