/*
 * radik_oracle.c -- CPU oracle for the radix top-k path.
 *
 * TEST INFRASTRUCTURE ONLY.  The product (libradik_hip.so and the host mirrors
 * above it) never links, loads or calls this file.  Only tests/, the smoke
 * check and bench.py's CPU-baseline leg use it, and only as the checker.
 *
 * Reference: WilliBee/RadiK.jl (Julia).  Citations are file:line in that tree.
 *
 * Two independent restatements of the same result live here:
 *
 *   orc_topk_schedule()  the reference's own algorithm, src/algorithm.jl:162-368:
 *                        12/12/8-bit radix passes over the order-preserving key
 *                        (pass 1 on the raw input, passes 2/3 on the compacted
 *                        candidates), select_bin after each pass, the final
 *                        filter against the k-th key with a residual boundary
 *                        count, the final sort and the len<k padding.
 *   orc_topk_exact()     the contract written as a plain total order: take the
 *                        k smallest (rank key, position) pairs, then sort them by
 *                        (output key, position).  Quickselect + qsort.
 *
 * The contract both implement (the reference leaves these points open or
 * racy; the choices are the ones the HIP path implements):
 *   key(x)    = bits(x) ^ (sign ? 0xFFFFFFFF : 0x80000000)   utils.jl:29-33
 *   scaled(x) = WITHSCALE && !isnan(x) ? fl32(x - s) : x,    utils.jl:62-71
 *               s = first element of the task, or 0 if it is NaN/Inf.
 *               (NaN keeps its own bits instead of the reference's
 *               floatmin/floatmax substitution, utils.jl:69, which the
 *               reference's histogram pass does not apply, count_bin.jl:111.)
 *   rank(x)   = LARGEST ? ~key(scaled(x)) : key(scaled(x)); smaller is better.
 *   selection = the min(k, n) smallest (rank, position) pairs: ties at the
 *               k-th value go to the lowest position (the reference picks them
 *               in atomic order, filter.jl:115-119).  Keys, not float
 *               compares, so -0 < +0 and subnormals/Inf are exact (the
 *               reference filter compares floats, filter.jl:111,115).
 *   order     = ascending (ASCEND ? key(x) : ~key(x), position) of the
 *               ORIGINAL value x (algorithm.jl:340-352 sorts val_out, which
 *               holds originals).
 *   outputs   = val_out[i] = x; idx_out[i] = WITHIDXIN ? idx_in[pos] : pos+1
 *               (1-based, filter.jl:76-86); slots i >= n hold 0.0f / -1
 *               (algorithm.jl:358-365).
 *
 * Parity pins (what checks each function; tests/test_oracle.py):
 *   orc_key, orc_bin_id          known answers of test/test_utils.jl:9-23 and
 *                                the IEEE-754 closed form over special values.
 *   orc_warp_scan                test/test_utils.jl:53-106 (cumsum / reverse
 *                                cumsum) and numpy.cumsum.
 *   orc_cross_warp_reduction     test/test_select_bin.jl:13-55 known answers.
 *   orc_find_bin_and_write       test/test_select_bin.jl:79-188 known answers.
 *   orc_select_bin_blocked       test/test_select_bin.jl:194-352 known answers
 *                                and agreement with orc_select_bin.
 *   orc_select_bin               same known answers; brute-force cumulative
 *                                definition in numpy.
 *   orc_count_bin_ex             test/test_count_bin.jl:12-61 ramps (numpy
 *                                histogram of the same closed-form bins).
 *   orc_filter                   test/test_filter.jl known answer sets.
 *   orc_topk_exact               numpy.lexsort (a library stable sort) and
 *                                brute force over all subsets for n <= 12.
 *   orc_topk_schedule            agreement with orc_topk_exact on random,
 *                                tie-heavy and special-value data.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ORC_PLANT selects a deliberate mistake for the mutation tests
 * (tests/test_oracle.py builds variants with -DORC_PLANT=n and checks that
 * the pins catch each one); 0 in every real build. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#define ORC_LARGEST   1u
#define ORC_ASCEND    2u
#define ORC_WITHSCALE 4u
#define ORC_WITHIDXIN 8u

/* ------------------------------------------------------------------------ */
/* Key transform and digits                                                  */
/* ------------------------------------------------------------------------ */

static inline uint32_t f2u(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static inline float u2f(uint32_t u) { float v; memcpy(&v, &u, 4); return v; }

/* utils.jl:29-33 with LEFT = RIGHT = 0: mask = ((~(bits >> 31)) + 1) | 0x80000000. */
uint32_t orc_key(float v)
{
    uint32_t bits = f2u(v);
    uint32_t mask = ((~(bits >> 31)) + 1u) | 0x80000000u;
    if (ORC_PLANT == 1) mask = 0x80000000u; /* planted: sign-magnitude order kept */
    return bits ^ mask;
}

/* Inverse of orc_key (the key is a bijection on 32-bit patterns). */
float orc_key_inv(uint32_t key)
{
    uint32_t bits = (key & 0x80000000u) ? (key ^ 0x80000000u) : ~key;
    return u2f(bits);
}

/* utils.jl:29-33: ((bits ^ mask) << LEFT) >>> RIGHT on UInt32. */
uint32_t orc_bin_id(float v, int left, int right)
{
    uint32_t k = orc_key(v);
    k = (left >= 32) ? 0u : (uint32_t)(k << left);
    return (right >= 32) ? 0u : (k >> right);
}

/* utils.jl:62-66. */
float orc_sample_scaler(float first, int withscale)
{
    if (!withscale) return 0.0f;
    if (ORC_PLANT == 4) return first; /* planted: Inf/NaN scaler not replaced */
    return (isnan(first) || isinf(first)) ? 0.0f : first;
}

/* Contract form of utils.jl:68-71 (NaN keeps its bits, see header). */
float orc_scaled(float v, float s, int withscale)
{
    if (!withscale || isnan(v)) return v;
    volatile float r = v - s; /* one IEEE binary32 subtraction, RNE */
    return r;
}

/* utils.jl:68-71 exactly as the reference writes it (NaN -> floatmin/floatmax). */
float orc_apply_scaling_ref(float v, float s, int withscale, int largest)
{
    float p = isnan(v) ? (largest ? 1.17549435e-38f : 3.40282347e+38f) : v;
    if (!withscale) return p;
    volatile float r = p - s;
    return r;
}

static inline uint32_t rank_key(float v, float s, uint32_t flags)
{
    uint32_t k = orc_key(orc_scaled(v, s, (flags & ORC_WITHSCALE) != 0));
    return (flags & ORC_LARGEST) ? ~k : k;
}

/* ------------------------------------------------------------------------ */
/* Scans (utils.jl:84-100, select_bin.jl:22-58)                              */
/* ------------------------------------------------------------------------ */

/* utils.jl:84-100: shuffle scan inside groups of ws lanes.  up != 0 gives the
 * inclusive prefix ("Up": lane > offset adds the lane offset below); up == 0
 * the inclusive suffix ("Down": lane + offset <= ws adds the lane above).
 * Simulated round by round so every lane reads the previous round's value. */
void orc_warp_scan(int64_t *v, int64_t n, int ws, int up)
{
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int off = 1; off < ws; off <<= 1) {
        memcpy(tmp, v, sizeof(int64_t) * (size_t)n);
        for (int64_t i = 0; i < n; i++) {
            int64_t lane = (i % ws) + 1; /* 1-based */
            int64_t base = i - (lane - 1);
            if (up) {
                if (lane > off) v[i] = tmp[base + lane - 1 - off] + tmp[i];
            } else {
                if (lane + off <= ws && base + lane - 1 + off < n)
                    v[i] = tmp[base + lane - 1 + off] + tmp[i];
            }
        }
    }
    free(tmp);
}

/* select_bin.jl:22-58: tree pass across warps.  prefix_sum holds the per-thread
 * warp-scanned sums (BLOCK entries); sums is each thread's own running value.
 * After the call sums[t] is the block-level inclusive scan (suffix for LARGEST). */
void orc_cross_warp_reduction(int64_t *prefix_sum, int64_t *sums, int block, int warp, int largest)
{
    int nw = block / warp;
    int64_t *nxt = (int64_t *)malloc(sizeof(int64_t) * (size_t)block);
    for (int i = 1; i < nw; i *= 2) {
        for (int t = 0; t < block; t++) {
            int w = t / warp;
            int64_t s = sums[t];
            if (largest) {
                if (w + i < nw) s += prefix_sum[(w + i) * warp];          /* first thread of warp w+i */
            } else {
                if (w - i >= 0) s += prefix_sum[(w - i + 1) * warp - 1];  /* last thread of warp w-i  */
            }
            nxt[t] = s;
        }
        for (int t = 0; t < block; t++) { sums[t] = nxt[t]; prefix_sum[t] = nxt[t]; }
    }
    free(nxt);
}

/* select_bin.jl:125-159.  counts: this thread's UNROLL bins; thread_x is
 * 1-based.  Writes (bin 1-based, residual k, bin count) into out[0..2] and
 * returns 1 when neighbor < old_k <= sum, else leaves out untouched. */
int orc_find_bin_and_write(const int64_t *counts, int64_t old_k, int64_t neighbor, int64_t sum,
                           int thread_x, int unroll, int largest, int64_t *out)
{
    if (!(neighbor < old_k && old_k <= sum)) return 0;
    old_k -= neighbor;
    for (int j = 0; j < unroll; j++) {
        int i = largest ? (unroll - j) : (j + 1); /* 1-based bin within the thread */
        if (counts[i - 1] >= old_k) {
            out[0] = (int64_t)(thread_x - 1) * unroll + i;
            out[1] = old_k;
            out[2] = counts[i - 1];
            return 1;
        }
        old_k -= counts[i - 1];
    }
    return 0;
}

/* select_bin.jl:268-348 simulated with its thread structure: load UNROLL bins
 * per thread, warp scan (Down for LARGEST, Up otherwise), cross-warp tree pass,
 * neighbour, find_bin_and_write.  Returns 1 and (bin 1-based, k', count) when
 * some thread's range holds k; 0 when k exceeds the total (nothing written,
 * select_bin.jl:139). */
int orc_select_bin_blocked(const int64_t *hist, int histlen, int block, int warp,
                           int64_t k, int largest, int64_t *out)
{
    int unroll = histlen / block;
    if (unroll * block != histlen || block % warp != 0) return -1;
    int64_t *sums = (int64_t *)calloc((size_t)block, sizeof(int64_t));
    int64_t *ps = (int64_t *)calloc((size_t)block, sizeof(int64_t));
    for (int t = 0; t < block; t++)
        for (int i = 0; i < unroll; i++) sums[t] += hist[t * unroll + i];
    orc_warp_scan(sums, block, warp, largest ? 0 : 1);
    memcpy(ps, sums, sizeof(int64_t) * (size_t)block);
    orc_cross_warp_reduction(ps, sums, block, warp, largest);
    int found = 0;
    for (int t = 0; t < block && !found; t++) {
        int64_t nb;
        if (largest) nb = (t + 1 < block) ? ps[t + 1] : 0;   /* compute_neighbor, select_bin.jl:89-99 */
        else         nb = (t > 0) ? ps[t - 1] : 0;
        found = orc_find_bin_and_write(hist + (size_t)t * unroll, k, nb, sums[t], t + 1, unroll, largest, out);
    }
    free(sums);
    free(ps);
    return found;
}

/* The bin rule of select_bin.jl:125-159 as a cumulative count: for LARGEST the
 * bins are walked from the top.  Returns 1 with (bin 1-based, k', count). */
int orc_select_bin(const int64_t *hist, int histlen, int64_t k, int largest, int64_t *out)
{
    int64_t before = 0;
    for (int j = 0; j < histlen; j++) {
        int b = largest ? (histlen - 1 - j) : j;
        if (before < k && k <= before + hist[b] + (ORC_PLANT == 5 ? 1 : 0)) { /* plant 5: off by one */
            out[0] = b + 1;
            out[1] = k - before;
            out[2] = hist[b];
            return 1;
        }
        before += hist[b];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Histograms                                                                */
/* ------------------------------------------------------------------------ */

/* count_bin.jl:110-116 / 119-187 for one task: bins of
 * get_bin_id(apply_scaling(v)) over the task, NaN skipped (count_bin.jl:111). */
void orc_count_bin_ex(const float *data, int64_t n, int left, int right, int histlen,
                      int withscale, int largest, int64_t *hist)
{
    memset(hist, 0, sizeof(int64_t) * (size_t)histlen);
    if (n <= 0) return;
    float s = orc_sample_scaler(data[0], withscale);
    for (int64_t i = 0; i < n; i++) {
        float v = data[i];
        if (isnan(v)) continue;
        uint32_t b = orc_bin_id(orc_apply_scaling_ref(v, s, withscale, largest), left, right);
        hist[b]++;
    }
}

/* Histogram of the contract's rank key: bins (rank >> shift) & (2^width - 1). */
void orc_rank_histogram(const float *data, int64_t n, uint32_t flags, int shift, int width, int64_t *hist)
{
    memset(hist, 0, sizeof(int64_t) * ((size_t)1 << width));
    if (n <= 0) return;
    float s = orc_sample_scaler(data[0], (flags & ORC_WITHSCALE) != 0);
    uint32_t m = (width >= 32) ? 0xFFFFFFFFu : ((1u << width) - 1u);
    for (int64_t i = 0; i < n; i++) hist[(rank_key(data[i], s, flags) >> shift) & m]++;
}

/* ------------------------------------------------------------------------ */
/* Final filter (filter.jl:106-139, 272-322), key form                       */
/* ------------------------------------------------------------------------ */

/* One task.  If n <= k every element is copied (filter.jl:272-322).  Else
 * admit rank < kth_rank, plus rank == kth_rank while the boundary count lasts
 * (filter.jl:115-119), taken in position order.  Writes up to k (value, pos)
 * pairs in position order; returns how many. */
int64_t orc_filter(const float *data, int64_t n, int64_t k, uint32_t kth_rank, int64_t boundary,
                   uint32_t flags, float *vals, int64_t *pos)
{
    int64_t c = 0;
    if (n <= k) {
        for (int64_t i = 0; i < n; i++) { vals[c] = data[i]; pos[c] = i; c++; }
        return c;
    }
    float s = orc_sample_scaler(data[0], (flags & ORC_WITHSCALE) != 0);
    for (int64_t i = 0; i < n; i++) {
        uint32_t r = rank_key(data[i], s, flags);
        int take = 0;
        if (r < kth_rank) take = 1;
        else if (r == kth_rank && boundary > 0) { boundary--; take = 1; }
        if (take && c < k) { vals[c] = data[i]; pos[c] = i; c++; }
    }
    return c;
}

/* ------------------------------------------------------------------------ */
/* Sorting helpers                                                           */
/* ------------------------------------------------------------------------ */

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Hoare-style quickselect: afterwards a[k-1] holds the k-th smallest and
 * a[0..k-1] the k smallest (all values distinct here). */
static void select_k_smallest(uint64_t *a, int64_t n, int64_t k)
{
    int64_t lo = 0, hi = n - 1, want = k - 1;
    uint64_t seed = 0x9E3779B97F4A7C15ull;
    while (lo < hi) {
        seed ^= seed << 13; seed ^= seed >> 7; seed ^= seed << 17;
        uint64_t pv = a[lo + (int64_t)(seed % (uint64_t)(hi - lo + 1))];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] < pv) i++;
            while (a[j] > pv) j--;
            if (i <= j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; i++; j--; }
        }
        if (want <= j) hi = j;
        else if (want >= i) lo = i;
        else break;
    }
}

/* Sorts m selected (value, pos) pairs by (ASCEND ? key : ~key, pos) and writes
 * the task's k output slots (padding beyond m). */
static void write_sorted(const float *data, const int64_t *sel_pos, int64_t m, int64_t k, uint32_t flags,
                         const void *idx_in, int idx_bytes, float *val_out, void *idx_out)
{
    uint64_t *c = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1));
    for (int64_t i = 0; i < m; i++) {
        uint32_t ok = orc_key(data[sel_pos[i]]);
        if (!(flags & ORC_ASCEND) != (ORC_PLANT == 3)) ok = ~ok; /* plant 3: order reversed */
        c[i] = ((uint64_t)ok << 32) | (uint64_t)(uint32_t)sel_pos[i];
    }
    qsort(c, (size_t)m, sizeof(uint64_t), cmp_u64);
    for (int64_t i = 0; i < k; i++) {
        float v = 0.0f;
        int64_t id = -1;
        if (i < m) {
            int64_t p = (int64_t)(uint32_t)(c[i] & 0xFFFFFFFFu);
            v = data[p];
            if (flags & ORC_WITHIDXIN)
                id = (idx_bytes == 8) ? ((const int64_t *)idx_in)[p] : (int64_t)((const int32_t *)idx_in)[p];
            else
                id = p + 1;
        }
        val_out[i] = v;
        if (idx_bytes == 8) ((int64_t *)idx_out)[i] = id;
        else ((int32_t *)idx_out)[i] = (int32_t)id;
    }
    free(c);
}

/* ------------------------------------------------------------------------ */
/* The contract as a direct order                                            */
/* ------------------------------------------------------------------------ */

static void topk_exact_task(const float *data, int64_t n, int64_t k, uint32_t flags,
                            const void *idx_in, int idx_bytes, float *val_out, void *idx_out)
{
    int64_t m = n < k ? n : k;
    uint64_t *c = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *sel = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    float s = (n > 0) ? orc_sample_scaler(data[0], (flags & ORC_WITHSCALE) != 0) : 0.0f;
    for (int64_t i = 0; i < n; i++) {
        uint64_t p = (ORC_PLANT == 2) ? (uint64_t)(n - 1 - i) : (uint64_t)i; /* planted: highest position wins ties */
        c[i] = ((uint64_t)rank_key(data[i], s, flags) << 32) | p;
    }
    if (m > 0 && m < n) select_k_smallest(c, n, m);
    for (int64_t i = 0; i < m; i++) {
        int64_t p = (int64_t)(uint32_t)(c[i] & 0xFFFFFFFFu);
        sel[i] = (ORC_PLANT == 2) ? (n - 1 - p) : p;
    }
    write_sorted(data, sel, m, k, flags, idx_in, idx_bytes, val_out, idx_out);
    free(c);
    free(sel);
}

/* Multi-task top-k over the concatenated input (task t starts at the sum of
 * the earlier lengths, algorithm.jl:206); outputs are [k, T] column-major.
 * idx_in (when WITHIDXIN) is indexed like val_in.  Returns 0 on success. */
int orc_topk_exact(const float *val_in, const void *idx_in, int idx_bytes, const int64_t *task_lens,
                   int64_t num_tasks, int64_t k, uint32_t flags, float *val_out, void *idx_out, int threads)
{
    if (k < 0 || num_tasks < 0 || (idx_bytes != 4 && idx_bytes != 8)) return 1;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(num_tasks + 1));
    off[0] = 0;
    for (int64_t t = 0; t < num_tasks; t++) off[t + 1] = off[t] + task_lens[t];
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t t = 0; t < num_tasks; t++) {
        const void *ii = (flags & ORC_WITHIDXIN) ? (const char *)idx_in + (size_t)off[t] * idx_bytes : NULL;
        topk_exact_task(val_in + off[t], task_lens[t], k, flags, ii, idx_bytes, val_out + (size_t)t * k,
                        (char *)idx_out + (size_t)t * k * idx_bytes);
    }
    (void)threads;
    free(off);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Other key types (beyond the reference, which takes Float32 only:          */
/* README.md:18-24, algorithm.jl:492-494): the same contract over the       */
/* order-preserving key of each type.                                        */
/*   kt 0 Float32  : orc_key (utils.jl:29-33)                                */
/*   kt 1 Float16, 2 BFloat16 (16-bit patterns h):                           */
/*        key16 = h ^ (sign ? 0xFFFF : 0x8000), the same closed form on 16   */
/*        bits; IEEE-754 totalOrder of the values                            */
/*   kt 3 Int32    : bits ^ 0x80000000 (two's complement order)              */
/* Values are copied as bit patterns; WITHSCALE is Float32-only.             */
/* ------------------------------------------------------------------------ */

uint32_t orc_key_typed(uint32_t bits, int kt)
{
    if (kt == 0) return orc_key(u2f(bits));
    if (kt == 3) return bits ^ 0x80000000u;
    {
        uint32_t h = bits & 0xFFFFu;
        return h ^ ((h & 0x8000u) ? 0xFFFFu : 0x8000u);
    }
}

static inline uint32_t load_bits(const void *p, int64_t i, int eb)
{
    if (eb == 2) return ((const uint16_t *)p)[i];
    return ((const uint32_t *)p)[i];
}

static void topk_exact_task_typed(int kt, const void *data, int64_t n, int64_t k, uint32_t flags,
                                  const void *idx_in, int idx_bytes, void *val_out, void *idx_out)
{
    const int eb = (kt == 1 || kt == 2) ? 2 : 4;
    const int64_t m = n < k ? n : k;
    const uint32_t xm = (flags & ORC_LARGEST) ? 0xFFFFFFFFu : 0u;
    const uint32_t am = (flags & ORC_ASCEND) ? 0u : 0xFFFFFFFFu;
    uint64_t *c = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++)
        c[i] = ((uint64_t)(orc_key_typed(load_bits(data, i, eb), kt) ^ xm) << 32) | (uint64_t)i;
    if (m > 0 && m < n) select_k_smallest(c, n, m);
    /* output order: (ASCEND ? key : ~key, position) */
    for (int64_t i = 0; i < m; i++) {
        const int64_t p = (int64_t)(uint32_t)c[i];
        c[i] = ((uint64_t)(orc_key_typed(load_bits(data, p, eb), kt) ^ am) << 32) | (uint64_t)p;
    }
    qsort(c, (size_t)m, sizeof(uint64_t), cmp_u64);
    for (int64_t i = 0; i < k; i++) {
        uint32_t v = 0;
        int64_t id = -1;
        if (i < m) {
            const int64_t p = (int64_t)(uint32_t)c[i];
            v = load_bits(data, p, eb);
            if (flags & ORC_WITHIDXIN)
                id = (idx_bytes == 8) ? ((const int64_t *)idx_in)[p] : (int64_t)((const int32_t *)idx_in)[p];
            else
                id = p + 1;
        }
        if (eb == 2) ((uint16_t *)val_out)[i] = (uint16_t)v;
        else ((uint32_t *)val_out)[i] = v;
        if (idx_bytes == 8) ((int64_t *)idx_out)[i] = id;
        else ((int32_t *)idx_out)[i] = (int32_t)id;
    }
    free(c);
}

int orc_topk_exact_typed(int kt, const void *val_in, const void *idx_in, int idx_bytes, const int64_t *task_lens,
                         int64_t num_tasks, int64_t k, uint32_t flags, void *val_out, void *idx_out, int threads)
{
    if (kt < 0 || kt > 3 || k < 0 || num_tasks < 0 || (idx_bytes != 4 && idx_bytes != 8)) return 1;
    if (kt != 0 && (flags & ORC_WITHSCALE)) return 1;
    const int eb = (kt == 1 || kt == 2) ? 2 : 4;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(num_tasks + 1));
    off[0] = 0;
    for (int64_t t = 0; t < num_tasks; t++) off[t + 1] = off[t] + task_lens[t];
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t t = 0; t < num_tasks; t++) {
        const void *ii = (flags & ORC_WITHIDXIN) ? (const char *)idx_in + (size_t)off[t] * idx_bytes : NULL;
        topk_exact_task_typed(kt, (const char *)val_in + (size_t)off[t] * eb, task_lens[t], k, flags, ii,
                              idx_bytes, (char *)val_out + (size_t)t * k * eb,
                              (char *)idx_out + (size_t)t * k * idx_bytes);
    }
    (void)threads;
    free(off);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference's schedule                                                  */
/* ------------------------------------------------------------------------ */

/* Per-task statistics of the schedule, for intermediate-level checks:
 * st[0..2] pass-1 (bin 1-based, k', count), st[3..5] pass 2, st[6..8] pass 3,
 * st[9] number of passes run, st[10] kth rank key. */
#define ORC_NSTAT 11

static void topk_schedule_task(const float *data, int64_t n, int64_t k, uint32_t flags,
                               const void *idx_in, int idx_bytes, float *val_out, void *idx_out, int64_t *st)
{
    int largest = (flags & ORC_LARGEST) != 0;
    int64_t m = n < k ? n : k;
    for (int i = 0; i < ORC_NSTAT; i++) st[i] = 0;
    int64_t *sel = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
    float *selv = (float *)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
    if (n <= k) {
        /* filter.jl:272-322: copy all. */
        int64_t c = orc_filter(data, n, k, 0, 0, flags, selv, sel);
        write_sorted(data, sel, c, k, flags, idx_in, idx_bytes, val_out, idx_out);
        free(sel); free(selv);
        return;
    }
    float s = orc_sample_scaler(data[0], (flags & ORC_WITHSCALE) != 0);
    /* Candidates are kept as key(scaled) values, as val_buffer_* hold scaled
     * floats (select_candidate.jl:162-165). */
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *next = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    int64_t hist[4096];
    int64_t res[3];
    int64_t kk = k;

    /* Pass 1, bits [31:20] (algorithm.jl:214-237): histogram, select_bin, compaction. */
    for (int b = 0; b < 4096; b++) hist[b] = 0;
    for (int64_t i = 0; i < n; i++) hist[orc_key(orc_scaled(data[i], s, (flags & ORC_WITHSCALE) != 0)) >> 20]++;
    orc_select_bin(hist, 4096, kk, largest, res);
    st[0] = res[0]; st[1] = res[1]; st[2] = res[2];
    kk = res[1];
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) {
        uint32_t key = orc_key(orc_scaled(data[i], s, (flags & ORC_WITHSCALE) != 0));
        if ((int64_t)(key >> 20) + 1 == res[0]) cand[nc++] = key; /* 1-based bin match, select_candidate.jl:109 */
    }
    int passes = 1;
    /* Pass 2, bits [19:8] (algorithm.jl:242-270) when more than one candidate. */
    int pass2 = 0;
    if (nc > 1) {
        for (int b = 0; b < 4096; b++) hist[b] = 0;
        for (int64_t i = 0; i < nc; i++) hist[(cand[i] >> 8) & 0xFFFu]++;
        orc_select_bin(hist, 4096, kk, largest, res);
        st[3] = res[0]; st[4] = res[1]; st[5] = res[2];
        kk = res[1];
        int64_t nn = 0;
        for (int64_t i = 0; i < nc; i++)
            if ((int64_t)((cand[i] >> 8) & 0xFFFu) + 1 == res[0]) next[nn++] = cand[i];
        uint32_t *t = cand; cand = next; next = t; nc = nn;
        pass2 = 1;
        passes = 2;
    }
    /* Pass 3, bits [7:0] (algorithm.jl:275-299) when pass 2 ran and more than one remains. */
    if (nc > 1 && pass2) {
        for (int b = 0; b < 256; b++) hist[b] = 0;
        for (int64_t i = 0; i < nc; i++) hist[cand[i] & 0xFFu]++;
        orc_select_bin(hist, 256, kk, largest, res);
        st[6] = res[0]; st[7] = res[1]; st[8] = res[2];
        kk = res[1];
        int64_t nn = 0;
        for (int64_t i = 0; i < nc; i++)
            if ((int64_t)(cand[i] & 0xFFu) + 1 == res[0]) next[nn++] = cand[i];
        uint32_t *t = cand; cand = next; next = t; nc = nn;
        passes = 3;
    }
    /* kth element = candidate #1 (algorithm.jl:306-308); residual k = boundary
     * count (algorithm.jl:325). */
    uint32_t kth = largest ? ~cand[0] : cand[0];
    st[9] = passes;
    st[10] = kth;
    int64_t c = orc_filter(data, n, k, kth, kk, flags, selv, sel);
    write_sorted(data, sel, c, k, flags, idx_in, idx_bytes, val_out, idx_out);
    free(cand); free(next); free(sel); free(selv);
}

/* The same schedule for one long task on several host threads (the CPU
 * baseline of a single task): the full-input steps -- pass-1 histogram,
 * pass-1 compaction, filter -- split the task into one contiguous range per
 * thread; the filter admits threshold-equal elements in position order
 * (each range's share of the boundary count from a prefix over the ranges),
 * so the outputs equal topk_schedule_task's. */
static void topk_schedule_task_par(const float *data, int64_t n, int64_t k, uint32_t flags,
                                   const void *idx_in, int idx_bytes, float *val_out, void *idx_out, int64_t *st,
                                   int nthr)
{
#ifndef _OPENMP
    topk_schedule_task(data, n, k, flags, idx_in, idx_bytes, val_out, idx_out, st);
    (void)nthr;
#else
    if (n <= k || nthr <= 1) {
        topk_schedule_task(data, n, k, flags, idx_in, idx_bytes, val_out, idx_out, st);
        return;
    }
    const int largest = (flags & ORC_LARGEST) != 0;
    const int ws = (flags & ORC_WITHSCALE) != 0;
    for (int i = 0; i < ORC_NSTAT; i++) st[i] = 0;
    const float s = orc_sample_scaler(data[0], ws);
    int64_t *hist = (int64_t *)calloc((size_t)nthr * 4096, sizeof(int64_t));
    int64_t *cnt = (int64_t *)calloc((size_t)nthr + 1, sizeof(int64_t));
    int64_t *cnt2 = (int64_t *)calloc((size_t)nthr + 1, sizeof(int64_t));
    int64_t res[3];
#pragma omp parallel num_threads(nthr)
    {
        const int t = omp_get_thread_num(), P = omp_get_num_threads();
        const int64_t a = n * t / P, b = n * (t + 1) / P;
        int64_t *h = hist + (size_t)t * 4096;
        for (int64_t i = a; i < b; i++) h[orc_key(orc_scaled(data[i], s, ws)) >> 20]++;
    }
    int64_t h1[4096];
    for (int b = 0; b < 4096; b++) {
        h1[b] = 0;
        for (int t = 0; t < nthr; t++) h1[b] += hist[(size_t)t * 4096 + b];
    }
    orc_select_bin(h1, 4096, k, largest, res);
    st[0] = res[0]; st[1] = res[1]; st[2] = res[2];
    int64_t kk = res[1];
    const int64_t bin1 = res[0];
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(res[2] > 0 ? res[2] : 1));
    uint32_t *next = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(res[2] > 0 ? res[2] : 1));
    /* pass-1 compaction: per-range counts, then each range writes at its offset */
#pragma omp parallel num_threads(nthr)
    {
        const int t = omp_get_thread_num(), P = omp_get_num_threads();
        const int64_t a = n * t / P, b = n * (t + 1) / P;
        int64_t c = 0;
        for (int64_t i = a; i < b; i++)
            if ((int64_t)(orc_key(orc_scaled(data[i], s, ws)) >> 20) + 1 == bin1) c++;
        cnt[t + 1] = c;
#pragma omp barrier
#pragma omp single
        for (int u = 0; u < P; u++) cnt[u + 1] += cnt[u];
        int64_t o = cnt[t];
        for (int64_t i = a; i < b; i++) {
            const uint32_t key = orc_key(orc_scaled(data[i], s, ws));
            if ((int64_t)(key >> 20) + 1 == bin1) cand[o++] = key;
        }
    }
    int64_t nc = cnt[nthr];
    int passes = 1, pass2 = 0;
    int64_t hh[4096];
    if (nc > 1) {
        for (int b = 0; b < 4096; b++) hh[b] = 0;
        for (int64_t i = 0; i < nc; i++) hh[(cand[i] >> 8) & 0xFFFu]++;
        orc_select_bin(hh, 4096, kk, largest, res);
        st[3] = res[0]; st[4] = res[1]; st[5] = res[2];
        kk = res[1];
        int64_t nn = 0;
        for (int64_t i = 0; i < nc; i++)
            if ((int64_t)((cand[i] >> 8) & 0xFFFu) + 1 == res[0]) next[nn++] = cand[i];
        uint32_t *tp = cand; cand = next; next = tp; nc = nn;
        pass2 = 1;
        passes = 2;
    }
    if (nc > 1 && pass2) {
        for (int b = 0; b < 256; b++) hh[b] = 0;
        for (int64_t i = 0; i < nc; i++) hh[cand[i] & 0xFFu]++;
        orc_select_bin(hh, 256, kk, largest, res);
        st[6] = res[0]; st[7] = res[1]; st[8] = res[2];
        kk = res[1];
        int64_t nn = 0;
        for (int64_t i = 0; i < nc; i++)
            if ((int64_t)(cand[i] & 0xFFu) + 1 == res[0]) next[nn++] = cand[i];
        uint32_t *tp = cand; cand = next; next = tp; nc = nn;
        passes = 3;
    }
    const uint32_t kth = largest ? ~cand[0] : cand[0];
    st[9] = passes;
    st[10] = kth;
    /* filter: strictly better elements, plus the first kk threshold-equal ones
     * in position order */
    int64_t *sel = (int64_t *)malloc(sizeof(int64_t) * (size_t)k);
    float *selv = (float *)malloc(sizeof(float) * (size_t)k);
    for (int t = 0; t <= nthr; t++) { cnt[t] = 0; cnt2[t] = 0; }
#pragma omp parallel num_threads(nthr)
    {
        const int t = omp_get_thread_num(), P = omp_get_num_threads();
        const int64_t a = n * t / P, b = n * (t + 1) / P;
        int64_t lt = 0, eq = 0;
        for (int64_t i = a; i < b; i++) {
            const uint32_t r = rank_key(data[i], s, flags);
            lt += r < kth;
            eq += r == kth;
        }
        cnt[t + 1] = lt;
        cnt2[t + 1] = eq;
#pragma omp barrier
#pragma omp single
        for (int u = 0; u < P; u++) {
            cnt2[u + 1] += cnt2[u];  /* equal elements before range u + 1 */
        }
        /* range t admits equal elements with global equal-rank < kk */
        const int64_t e0 = cnt2[t];
        int64_t admit = kk - e0;
        if (admit < 0) admit = 0;
        if (admit > eq) admit = eq;
        cnt[t + 1] += admit;
#pragma omp barrier
#pragma omp single
        for (int u = 0; u < P; u++) cnt[u + 1] += cnt[u];
        int64_t o = cnt[t], left = admit;
        for (int64_t i = a; i < b; i++) {
            const uint32_t r = rank_key(data[i], s, flags);
            int take = r < kth;
            if (!take && r == kth && left > 0) { left--; take = 1; }
            if (take && o < k) { selv[o] = data[i]; sel[o] = i; o++; }
        }
    }
    int64_t c = cnt[nthr] < k ? cnt[nthr] : k;
    write_sorted(data, sel, c, k, flags, idx_in, idx_bytes, val_out, idx_out);
    free(sel); free(selv); free(cand); free(next); free(hist); free(cnt); free(cnt2);
#endif
}

int orc_topk_schedule(const float *val_in, const void *idx_in, int idx_bytes, const int64_t *task_lens,
                      int64_t num_tasks, int64_t k, uint32_t flags, float *val_out, void *idx_out,
                      int64_t *stats, int threads)
{
    if (k < 0 || num_tasks < 0 || (idx_bytes != 4 && idx_bytes != 8)) return 1;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(num_tasks + 1));
    off[0] = 0;
    for (int64_t t = 0; t < num_tasks; t++) off[t + 1] = off[t] + task_lens[t];
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
    if (num_tasks == 1 && threads > 1) {
        /* one task: the threads split its element range */
        int64_t local[ORC_NSTAT];
        const void *ii = (flags & ORC_WITHIDXIN) ? idx_in : NULL;
        topk_schedule_task_par(val_in, task_lens[0], k, flags, ii, idx_bytes, val_out, idx_out, local, threads);
        if (stats) memcpy(stats, local, sizeof(local));
        free(off);
        return 0;
    }
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t t = 0; t < num_tasks; t++) {
        int64_t local[ORC_NSTAT];
        const void *ii = (flags & ORC_WITHIDXIN) ? (const char *)idx_in + (size_t)off[t] * idx_bytes : NULL;
        topk_schedule_task(val_in + off[t], task_lens[t], k, flags, ii, idx_bytes, val_out + (size_t)t * k,
                           (char *)idx_out + (size_t)t * k * idx_bytes, local);
        if (stats) memcpy(stats + (size_t)t * ORC_NSTAT, local, sizeof(local));
    }
    (void)threads;
    free(off);
    return 0;
}

int orc_nstat(void) { return ORC_NSTAT; }

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
