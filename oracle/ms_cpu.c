/*
 * ms_cpu.c -- the CPU baseline of the fused select + agg_sum: a streaming
 * restatement of the reference's decompress_and_process_batch with the
 * select core feeding the agg_sum core (format.h:164-195,
 * select_compr.h:80-89, agg_sum_compr.h:57-153) for every layout of the
 * format cascade sweep, block by block, never materialising the column.
 *
 * MEASUREMENT INFRASTRUCTURE ONLY: timed by bench.py's cpu_baseline leg
 * and checked by tests/ against the oracle; the product never loads it.
 *
 * Two processing styles, as the reference's vector extensions:
 *   simd = 0  scalar: value by value (the reference's vectorlib::scalar);
 *   simd = 1  AVX-512: one 512-bit register holds the 8 interleaved lanes
 *             of an 8-lane vertical bit-packed block, so a row of 8 codes
 *             is cut with one shift / or / and per word at offsets fixed
 *             by the block's bit width (the generated unpack routines of
 *             vbp_routine_gen.py:182-243, specialised per bit width by a
 *             switch as vbp_commons.h:77-377 does), compared with an
 *             unsigned mask compare and added under that mask
 *             (select_compr.h:80-89, agg_sum_compr.h:57-71).  Built with
 *             target attributes and used only where the CPU has AVX-512F.
 * Threads split a column's blocks (pages for for_f, runs for rle); delta_f
 * with more than one thread first sums each thread's range of deltas to
 * find its starting carry (the reference has no threading).
 */
#include "ms_oracle.h"

#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define U64MAX 0xFFFFFFFFFFFFFFFFull

/* predicate as [lo, hi] (optionally negated), as the GPU path states it */
typedef struct cpred {
    uint64_t lo, span;
    int neg, empty;
} cpred;

static cpred make_cpred(uint32_t cmp, uint64_t v)
{
    cpred p = {0, U64MAX, 0, 0};
    uint64_t hi = U64MAX;
    switch (cmp) {
    case MS_CMP_EQ: p.lo = v; hi = v; break;
    case MS_CMP_NE: p.lo = v; hi = v; p.neg = 1; break;
    case MS_CMP_LT: if (v == 0) p.empty = 1; else hi = v - 1; break;
    case MS_CMP_LE: hi = v; break;
    case MS_CMP_GT: if (v == U64MAX) p.empty = 1; else p.lo = v + 1; break;
    case MS_CMP_GE: p.lo = v; break;
    default: p.empty = 1; break;
    }
    p.span = hi - p.lo;
    return p;
}

static inline int ptake(const cpred *p, uint64_t x)
{
    const int in = !p->empty && (x - p->lo) <= p->span;
    return in != p->neg;
}

/* ------------------------------------------------------------------ */
/* Block walk: every 8-lane vbp block of 512 values of the compressed   */
/* part, with its payload offset, bit width and the transform its codes */
/* take (FOR reference; delta blocks are marked).                      */
/* ------------------------------------------------------------------ */
typedef struct blk {
    uint64_t off;   /* payload byte offset in the image */
    uint64_t ref;   /* for_f reference value (0 otherwise) */
    uint32_t bw;
} blk;

/* dynamic_vbp_f pages (dynamic_vbp.h:188-257): P meta bytes, then the
 * blocks; 0xFF marks a missing block of the call's last page. */
static uint64_t walk_dvbp(const uint8_t *img, uint64_t pos, uint64_t n, uint64_t B, uint64_t P, uint64_t ref,
                          blk *out, uint64_t *nout)
{
    uint64_t done = 0;
    while (done < n) {
        const uint8_t *meta = img + pos;
        pos += P;
        for (uint64_t q = 0; q < P && done < n; q++) {
            if (meta[q] == MS_VBP_BW_NOBLOCK) continue;
            for (uint64_t s = 0; s < B / 512; s++) {
                blk *b = &out[(*nout)++];
                b->off = pos + s * 512 * meta[q] / 8;
                b->ref = ref;
                b->bw = meta[q];
            }
            pos += B * meta[q] / 8;
            done += B;
        }
    }
    return pos;
}

/* All blocks of the compressed part of an 8-lane layout this baseline
 * streams: dynamic_vbp_f<512k, P, 8>, delta_f<L, 8, dynamic_vbp_f<512k,
 * P', 8>> and for_f<L, P, dynamic_vbp_f<512k, P', 8>>. */
static int walk(const oc_col *in, uint64_t nc, blk **blocks, uint64_t *nb)
{
    const ms_format_desc *f = &in->f;
    blk *out = (blk *)malloc(sizeof(blk) * (nc / 512 + 1));
    uint64_t cnt = 0;
    if (!out) return MS_ERR_NOMEM;
    if (f->kind == MS_DYNAMIC_VBP) {
        walk_dvbp(in->data, 0, nc, f->block, f->page, 0, out, &cnt);
    } else if (f->kind == MS_DELTA) {
        uint64_t pos = 0;
        /* one inner call per block (delta.h:106-156); the lanes' running
         * sums continue across blocks, so any block boundary splits work */
        for (uint64_t o = 0; o < nc; o += f->block)
            pos = walk_dvbp(in->data, pos, f->block, f->inner_block, f->inner_page, 0, out, &cnt);
    } else {   /* MS_FOR (for.h:195-272): pages of P references, then the inner calls */
        uint64_t pos = 0, done = 0;
        while (done < nc) {
            const uint64_t *refs = (const uint64_t *)(in->data + pos);
            pos += 8ull * f->page;
            for (uint64_t q = 0; q < f->page && done < nc; q++) {
                if (refs[q] == U64MAX) break;
                pos = walk_dvbp(in->data, pos, f->block, f->inner_block, f->inner_page, refs[q], out, &cnt);
                done += f->block;
            }
        }
    }
    *blocks = out;
    *nb = cnt;
    return MS_OK;
}

static int streamable(const ms_format_desc *f)
{
    if (f->kind == MS_UNCOMPR || f->kind == MS_RLE) return 1;
    if (f->kind == MS_DYNAMIC_VBP) return f->step == 8 && f->block % 512 == 0;
    if (f->kind == MS_DELTA) return f->step == 8 && f->inner_kind == MS_DYNAMIC_VBP && f->inner_step == 8 &&
                                    f->inner_block % 512 == 0 && f->block % f->inner_block == 0;
    if (f->kind == MS_FOR) return f->inner_kind == MS_DYNAMIC_VBP && f->inner_step == 8 &&
                                  f->inner_block % 512 == 0 && f->block % f->inner_block == 0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Scalar style: one value at a time (vbp_l unpack, vbp_format_rndaccess */
/* .h:60-87 with S = 8).                                               */
/* ------------------------------------------------------------------ */
static uint64_t blk_scalar(const uint64_t *w, unsigned bw, uint64_t ref, int delta, uint64_t *prev, const cpred *p)
{
    const uint64_t mask = bw >= 64 ? U64MAX : ((1ull << bw) - 1ull);
    uint64_t s = 0;
    for (unsigned k = 0; k < 64; k++) {
        const unsigned b = k * bw, wi = b >> 6, off = b & 63;
        for (unsigned j = 0; j < 8; j++) {
            uint64_t v = w[8 * wi + j] >> off;
            if (off + bw > 64) v |= w[8 * (wi + 1) + j] << (64 - off);
            v &= mask;
            uint64_t x;
            if (delta) x = prev[j] += v;
            else x = v + ref;
            if (ptake(p, x)) s += x;
        }
    }
    return s;
}

/* ------------------------------------------------------------------ */
/* AVX-512 style                                                        */
/* ------------------------------------------------------------------ */
#define AVX512 __attribute__((target("avx512f,avx512dq,avx512vl,avx512bw")))

static inline AVX512 __attribute__((always_inline)) __m512i blk_avx512_bw(const uint64_t *w, const unsigned bw,
                                                                           __m512i ref, const int delta,
                                                                           __m512i *prev, __m512i lo, __m512i span,
                                                                           const int neg, __m512i acc)
{
    const __m512i mask = _mm512_set1_epi64((long long)(bw >= 64 ? U64MAX : ((1ull << bw) - 1ull)));
    __m512i pv = *prev;
#pragma GCC unroll 64
    for (unsigned k = 0; k < 64; k++) {
        const unsigned b = k * bw, wi = b >> 6, off = b & 63;
        __m512i v = _mm512_srli_epi64(_mm512_loadu_si512((const void *)(w + 8 * wi)), off);
        if (off + bw > 64)
            v = _mm512_or_si512(v, _mm512_slli_epi64(_mm512_loadu_si512((const void *)(w + 8 * (wi + 1))), 64 - off));
        if (bw < 64) v = _mm512_and_si512(v, mask);
        __m512i x;
        if (delta) x = pv = _mm512_add_epi64(pv, v);
        else x = _mm512_add_epi64(v, ref);
        __mmask8 m = _mm512_cmple_epu64_mask(_mm512_sub_epi64(x, lo), span);
        if (neg) m = (__mmask8)~m;
        acc = _mm512_mask_add_epi64(acc, m, acc, x);
    }
    *prev = pv;
    return acc;
}

#define MS_CPU_BW_CASES(X)                                                                                         \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)    \
    X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39) \
    X(40) X(41) X(42) X(43) X(44) X(45) X(46) X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58) \
    X(59) X(60) X(61) X(62) X(63) X(64)

static AVX512 __attribute__((noinline)) __m512i blk_avx512(const uint64_t *w, unsigned bw, __m512i ref, int delta,
                                                            __m512i *prev, __m512i lo, __m512i span, int neg,
                                                            __m512i acc)
{
    switch (bw) {
#define MS_CASE(B)                                                                                             \
    case B:                                                                                                    \
        return delta ? (neg ? blk_avx512_bw(w, B, ref, 1, prev, lo, span, 1, acc)                              \
                            : blk_avx512_bw(w, B, ref, 1, prev, lo, span, 0, acc))                             \
                     : (neg ? blk_avx512_bw(w, B, ref, 0, prev, lo, span, 1, acc)                              \
                            : blk_avx512_bw(w, B, ref, 0, prev, lo, span, 0, acc));
        MS_CPU_BW_CASES(MS_CASE)
#undef MS_CASE
    default: return acc;
    }
}

static AVX512 uint64_t blocks_avx512(const oc_col *in, const blk *bl, uint64_t b0, uint64_t b1, int delta,
                                     uint64_t *carry, const cpred *p)
{
    __m512i acc = _mm512_setzero_si512();
    __m512i prev = _mm512_loadu_si512((const void *)carry);
    const __m512i lo = _mm512_set1_epi64((long long)p->lo), span = _mm512_set1_epi64((long long)p->span);
    if (p->empty && !p->neg) return 0;
    for (uint64_t b = b0; b < b1; b++) {
        const uint64_t *w = (const uint64_t *)(in->data + bl[b].off);
        if (p->empty)   /* negated empty range: every value */
            acc = blk_avx512(w, bl[b].bw, _mm512_set1_epi64((long long)bl[b].ref), delta, &prev,
                             _mm512_setzero_si512(), _mm512_set1_epi64(-1), 0, acc);
        else
            acc = blk_avx512(w, bl[b].bw, _mm512_set1_epi64((long long)bl[b].ref), delta, &prev, lo, span,
                             p->neg, acc);
    }
    _mm512_storeu_si512((void *)carry, prev);
    return (uint64_t)_mm512_reduce_add_epi64(acc);
}

/* rle_f: a run passes as a whole; AVX-512 takes 8 runs per step. */
static AVX512 uint64_t rle_avx512(const uint64_t *pairs, uint64_t r0, uint64_t r1, const cpred *p)
{
    __m512i acc = _mm512_setzero_si512();
    const __m512i lo = _mm512_set1_epi64((long long)p->lo), span = _mm512_set1_epi64((long long)p->span);
    const __m512i iv = _mm512_set_epi64(14, 12, 10, 8, 6, 4, 2, 0), il = _mm512_set_epi64(15, 13, 11, 9, 7, 5, 3, 1);
    uint64_t r = r0, s = 0;
    for (; r + 8 <= r1; r += 8) {
        const __m512i a = _mm512_loadu_si512((const void *)(pairs + 2 * r));
        const __m512i b = _mm512_loadu_si512((const void *)(pairs + 2 * r + 8));
        const __m512i v = _mm512_permutex2var_epi64(a, iv, b), l = _mm512_permutex2var_epi64(a, il, b);
        __mmask8 m = p->empty ? 0 : _mm512_cmple_epu64_mask(_mm512_sub_epi64(v, lo), span);
        if (p->neg) m = (__mmask8)~m;
        acc = _mm512_mask_add_epi64(acc, m, acc, _mm512_mullo_epi64(v, l));
    }
    for (; r < r1; r++)
        if (ptake(p, pairs[2 * r])) s += pairs[2 * r] * pairs[2 * r + 1];
    return s + (uint64_t)_mm512_reduce_add_epi64(acc);
}

static uint64_t blocks_scalar(const oc_col *in, const blk *bl, uint64_t b0, uint64_t b1, int delta, uint64_t *carry,
                              const cpred *p)
{
    uint64_t s = 0;
    for (uint64_t b = b0; b < b1; b++)
        s += blk_scalar((const uint64_t *)(in->data + bl[b].off), bl[b].bw, bl[b].ref, delta, carry, p);
    return s;
}

static uint64_t rle_scalar(const uint64_t *pairs, uint64_t r0, uint64_t r1, const cpred *p)
{
    uint64_t s = 0;
    for (uint64_t r = r0; r < r1; r++)
        if (ptake(p, pairs[2 * r])) s += pairs[2 * r] * pairs[2 * r + 1];
    return s;
}

static uint64_t raw_sum(const uint64_t *x, uint64_t n, const cpred *p)
{
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; i++)
        if (ptake(p, x[i])) s += x[i];
    return s;
}

/* ------------------------------------------------------------------ */
/* Threads                                                              */
/* ------------------------------------------------------------------ */
typedef struct job {
    const oc_col *in;
    const blk *bl;
    uint64_t b0, b1;
    int simd, delta, pass;   /* pass 0 (delta only): lane totals; pass 1: select + sum */
    cpred p;
    uint64_t carry[8];
    uint64_t sum;
} job;

/* Lane totals of a range of delta blocks (the starting carry of the next
 * thread's range): the same decode under a predicate that takes only 0. */
static void delta_totals(const oc_col *in, const blk *bl, uint64_t b0, uint64_t b1, uint64_t *tot, int simd)
{
    const cpred zero = {0, 0, 0, 0};
    memset(tot, 0, 64);
    if (simd) (void)blocks_avx512(in, bl, b0, b1, 1, tot, &zero);
    else (void)blocks_scalar(in, bl, b0, b1, 1, tot, &zero);
}

static void *worker(void *arg)
{
    job *jb = (job *)arg;
    const ms_format_desc *f = &jb->in->f;
    if (jb->pass == 0) {
        delta_totals(jb->in, jb->bl, jb->b0, jb->b1, jb->carry, jb->simd);
        return NULL;
    }
    if (f->kind == MS_RLE) {
        const uint64_t *pairs = (const uint64_t *)jb->in->data;
        jb->sum = jb->simd ? rle_avx512(pairs, jb->b0, jb->b1, &jb->p) : rle_scalar(pairs, jb->b0, jb->b1, &jb->p);
    } else if (f->kind == MS_UNCOMPR) {
        jb->sum = raw_sum((const uint64_t *)jb->in->data + jb->b0, jb->b1 - jb->b0, &jb->p);
    } else {
        jb->sum = jb->simd ? blocks_avx512(jb->in, jb->bl, jb->b0, jb->b1, jb->delta, jb->carry, &jb->p)
                           : blocks_scalar(jb->in, jb->bl, jb->b0, jb->b1, jb->delta, jb->carry, &jb->p);
    }
    return NULL;
}

static void run_jobs(job *jobs, int threads)
{
    pthread_t th[256];
    if (threads == 1) {
        worker(&jobs[0]);
        return;
    }
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &jobs[t]);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
}

int oc_cpu_has_avx512(void)
{
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq") &&
           __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512bw");
}

int oc_cpu_select_sum(uint64_t *result, uint32_t cmp, uint64_t val, const oc_col *in, int threads, int simd)
{
    const ms_format_desc *f = &in->f;
    const cpred p = make_cpred(cmp, val);
    job jobs[256];
    blk *bl = NULL;
    uint64_t units = 0, rest, nc, r, s = 0;
    int delta = f->kind == MS_DELTA;
    if (!streamable(f)) return MS_ERR_INVALID;
    if (simd && !oc_cpu_has_avx512()) return MS_ERR_INVALID;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    rest = oc_count_values_uncompr(in);
    nc = in->count - rest;
    if (f->kind == MS_RLE) {
        units = in->size_used / 16;
        rest = 0;
    } else if (f->kind == MS_UNCOMPR) {
        units = in->count;
        rest = 0;
    } else {
        int rc = walk(in, nc, &bl, &units);
        if (rc) return rc;
    }
    memset(jobs, 0, sizeof(job) * (size_t)threads);
    for (int t = 0; t < threads; t++) {
        const uint64_t b0 = units * (uint64_t)t / (uint64_t)threads;
        const uint64_t b1 = units * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t].in = in;
        jobs[t].bl = bl;
        jobs[t].b0 = b0;
        jobs[t].b1 = b1;
        jobs[t].simd = simd;
        jobs[t].delta = delta;
        jobs[t].p = p;
        jobs[t].pass = 0;
    }
    if (delta && threads > 1) {
        uint64_t run[8] = {0};
        run_jobs(jobs, threads);
        for (int t = 0; t < threads; t++) {
            uint64_t tot[8];
            memcpy(tot, jobs[t].carry, 64);
            memcpy(jobs[t].carry, run, 64);
            for (int j = 0; j < 8; j++) run[j] += tot[j];
        }
    } else {
        for (int t = 0; t < threads; t++) memset(jobs[t].carry, 0, 64);
    }
    for (int t = 0; t < threads; t++) jobs[t].pass = 1;
    run_jobs(jobs, threads);
    for (int t = 0; t < threads; t++) s += jobs[t].sum;
    if (rest) {
        oc_rest_offset(f, in->size_compr, &r);
        s += raw_sum((const uint64_t *)(in->data + r), rest, &p);
    }
    free(bl);
    *result = s;
    return MS_OK;
}
