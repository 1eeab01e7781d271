/*
 * ms_oracle.c -- CPU restatement of MorphStore's compression-aware path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
 * product.  Each function names the reference code it restates; paths are
 * relative to the reference root.  All arithmetic is uint64_t mod 2^64.
 */
#include "ms_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define U64MAX 0xFFFFFFFFFFFFFFFFull

/* ------------------------------------------------------------------ */
/* Helpers: math.h:60-153, alignment_helper.h:25-46                    */
/* ------------------------------------------------------------------ */

/* math.h:108-113: effective bit width, 0 has one effective bit. */
unsigned oc_effective_bitwidth(uint64_t v) { return 64u - (unsigned)__builtin_clzll(v | 1u); }

/* math.h:146-153 */
static uint64_t bwi_single_max(uint64_t B);
static uint64_t bw_max(unsigned bw) { return bw >= 64 ? U64MAX : ((1ull << bw) - 1ull); }

/* create_aligned_ptr with 64-byte alignment, relative to an aligned base. */
static uint64_t align64(uint64_t x) { return (x + 63ull) & ~63ull; }

static uint64_t round_down(uint64_t x, uint64_t f) { return x / f * f; }
static uint64_t round_up(uint64_t x, uint64_t f) { return (x + f - 1) / f * f; }

static void inner_of(const ms_format_desc *f, ms_format_desc *in)
{
    memset(in, 0, sizeof *in);
    in->kind = f->inner_kind;
    in->bw = f->inner_bw;
    in->block = f->inner_block;
    in->page = f->inner_page;
    in->step = f->inner_step;
}

/* Lane count of for_f's vector extension: it decides the reference-value
 * minimum (for.h:121-161) and packs the inner format. */
static unsigned for_lanes(const ms_format_desc *f)
{
    if (f->step) return f->step;
    if (f->inner_kind != MS_UNCOMPR && f->inner_step) return f->inner_step;
    return 1;
}

static int valid_leaf(const ms_format_desc *f)
{
    switch (f->kind) {
    case MS_UNCOMPR:
        return 1;
    case MS_STATIC_VBP:
        return f->bw >= 1 && f->bw <= 64 && f->step >= 1 && f->step <= 64;
    case MS_DYNAMIC_VBP:
        /* dynamic_vbp.h:45-53, 85-100 */
        return f->step >= 1 && f->step <= 64 && f->block > 0 && f->page > 0 &&
               f->block % (64u * f->step) == 0 && f->page % (8u * f->step) == 0;
    default:
        return 0;
    }
}

static int valid(const ms_format_desc *f)
{
    ms_format_desc in;
    uint64_t ibs;
    switch (f->kind) {
    case MS_UNCOMPR:
    case MS_STATIC_VBP:
    case MS_DYNAMIC_VBP:
        return valid_leaf(f);
    case MS_RLE:
        return 1;
    case MS_STATIC_VBP_PAD:   /* vbp_padding.h:45-66 */
        return f->bw >= 1 && f->bw <= 64 && f->step >= 1 && f->step <= 64;
    case MS_K_WISE_NS:        /* k_wise_ns.h: the SSE implementation, 2 values per block */
        return f->block == 2;
    case MS_GROUP_SIMPLE:     /* group_simple.h:64-131 with uint64_t, 8 E-byte alignment */
        return f->step >= 1 && f->step <= 64;
    case MS_BLOCKWISE_INDIVIDUAL: {   /* blockwise_individual.h:125-139 */
        ms_format_desc d;
        return f->block > 0 && f->block % 1024 == 0 && f->block <= (1u << 20) && f->step == 8 &&
               f->bw <= MS_BWI_SEL_STATIC && (f->bw != MS_BWI_SEL_STATIC || oc_bwi_inner(f->page, &d) == MS_OK);
    }
    case MS_DELTA:
    case MS_FOR:
        inner_of(f, &in);
        if (!valid_leaf(&in) || f->block == 0) return 0;
        if (oc_block_size(&in, &ibs) != MS_OK) return 0;
        if (f->block % ibs) return 0;            /* delta.h:80-84, for.h:86-90 */
        if (f->kind == MS_DELTA)
            return f->step >= 1 && f->step <= 64 && f->block % f->step == 0 &&
                   (in.kind == MS_UNCOMPR || in.step == f->step);
        return f->page > 0 && f->page % for_lanes(f) == 0 && f->block % for_lanes(f) == 0 &&
               (in.kind == MS_UNCOMPR || in.step == for_lanes(f));   /* for.h:50-55 */
    default:
        return 0;
    }
}

/* ------------------------------------------------------------------ */
/* Sizes: format.h:90-151, static_vbp.h:63-67, dynamic_vbp.h:106-120,  */
/* delta.h:92-104, for.h:100-112, rle.h:44-46, column.h:182-196        */
/* ------------------------------------------------------------------ */

int oc_block_size(const ms_format_desc *f, uint64_t *out)
{
    switch (f->kind) {
    case MS_UNCOMPR: *out = 1; return MS_OK;
    case MS_STATIC_VBP: *out = 64ull * f->step; return MS_OK;   /* vbp_format_rndaccess.h:76 */
    case MS_DYNAMIC_VBP:
    case MS_DELTA:
    case MS_FOR: *out = f->block; return MS_OK;
    case MS_STATIC_VBP_PAD: *out = (uint64_t)(64 / f->bw) * f->step; return MS_OK;   /* vbp_padding.h:64-66 */
    case MS_K_WISE_NS: *out = f->block; return MS_OK;
    case MS_GROUP_SIMPLE: *out = f->step; return MS_OK;                              /* t_GroupSizeLog */
    case MS_BLOCKWISE_INDIVIDUAL: *out = f->block; return MS_OK;                     /* blockwise_individual.h:155 */
    default: return MS_ERR_INVALID;                              /* rle_f has no m_BlockSize */
    }
}

int oc_size_max_byte(const ms_format_desc *f, uint64_t n, uint64_t *out)
{
    ms_format_desc in;
    uint64_t x;
    if (!valid(f)) return MS_ERR_INVALID;
    switch (f->kind) {
    case MS_UNCOMPR: *out = 8 * n; return MS_OK;
    case MS_STATIC_VBP: *out = n * f->bw / 8; return MS_OK;
    case MS_DYNAMIC_VBP: *out = 8 * n + (n / f->block) * f->page; return MS_OK;
    case MS_DELTA:
        inner_of(f, &in);
        return oc_size_max_byte(&in, n, out);
    case MS_FOR: {
        uint64_t pages = (n + (uint64_t)f->page * f->block - 1) / ((uint64_t)f->page * f->block);
        inner_of(f, &in);
        if (oc_size_max_byte(&in, f->block, &x) != MS_OK) return MS_ERR_INVALID;
        *out = pages * 8ull * f->page + (n / f->block) * x;
        return MS_OK;
    }
    case MS_RLE: *out = 16 * n; return MS_OK;
    case MS_STATIC_VBP_PAD:   /* vbp_padding.h:60-63 */
        *out = n / ((uint64_t)(64 / f->bw) * f->step) * f->step * 8 * 10;
        return MS_OK;
    case MS_K_WISE_NS: *out = 8 * n + n / f->block; return MS_OK;   /* k_wise_ns.h:126-130 */
    case MS_GROUP_SIMPLE: {   /* group_simple.h:93-121 (the fixed 1024-value call bound) */
        const uint64_t calls = (n + 1023) / 1024, sels = n / f->step / 2;
        *out = calls * 9 + round_up(sels, 2) + calls * (8ull * f->step - 1) + 8 * n;
        return MS_OK;
    }
    case MS_BLOCKWISE_INDIVIDUAL:   /* blockwise_individual.h:141-153 */
        *out = n / f->block * (bwi_single_max(f->block) + 64);
        return MS_OK;
    default: return MS_ERR_INVALID;
    }
}

int oc_size_max_byte_any_len(const ms_format_desc *f, uint64_t n, uint64_t *out)
{
    uint64_t bs, c;
    if (!valid(f)) return MS_ERR_INVALID;
    if (f->kind == MS_UNCOMPR) { *out = 8 * n; return MS_OK; }
    if (f->kind == MS_RLE) { *out = 16 * n; return MS_OK; }   /* rle.h:170-172 */
    oc_block_size(f, &bs);
    oc_size_max_byte(f, round_up(n, bs) + bs, &c);
    *out = c + 8 * (n % bs + bs) + 63;                          /* + get_size_with_alignment_padding */
    return MS_OK;
}

int oc_rest_offset(const ms_format_desc *f, uint64_t compr_end, uint64_t *out)
{
    uint64_t bs, m;
    if (f->kind == MS_UNCOMPR) { *out = align64(compr_end); return MS_OK; }
    if (oc_block_size(f, &bs) != MS_OK) return MS_ERR_INVALID;
    oc_size_max_byte(f, bs, &m);
    *out = align64(compr_end + 2 * m);
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* vbp_l<bw, S> (vbp_format_rndaccess.h:60-87; the routines emitted by */
/* vbp_routine_gen.py:61-180).  Value i is code k = i / S of lane      */
/* j = i % S; lane j's 64-bit word w sits at word index w*S + j.  Pack  */
/* ORs codes in unmasked (vbp_routine_gen.py:101,111); unpack masks.   */
/* ------------------------------------------------------------------ */

static void vbp_pack(const uint64_t *in, uint64_t n, unsigned bw, unsigned S, uint64_t *out)
{
    const uint64_t per_lane = n / S;
    for (unsigned j = 0; j < S; j++) {
        for (uint64_t k = 0; k < per_lane; k++) {
            const uint64_t v = in[k * S + j];
            const uint64_t b = k * bw;
            const uint64_t w = b >> 6;
            const unsigned off = (unsigned)(b & 63);
            out[w * S + j] |= v << off;
            if (off + bw > 64) out[(w + 1) * S + j] |= v >> (64 - off);
        }
    }
}

static void vbp_unpack(const uint64_t *in, uint64_t n, unsigned bw, unsigned S, uint64_t *out)
{
    const uint64_t per_lane = n / S;
    const uint64_t mask = bw_max(bw);
    for (unsigned j = 0; j < S; j++) {
        for (uint64_t k = 0; k < per_lane; k++) {
            const uint64_t b = k * bw;
            const uint64_t w = b >> 6;
            const unsigned off = (unsigned)(b & 63);
            uint64_t v = in[w * S + j] >> off;
            if (off + bw > 64) v |= in[(w + 1) * S + j] << (64 - off);
            out[k * S + j] = v & mask;
        }
    }
}

/* ------------------------------------------------------------------ */
/* dynamic_vbp_f (dynamic_vbp.h:122-308)                                */
/* ------------------------------------------------------------------ */

/* determine_max_bitwidth (dynamic_vbp.h:122-158): bw of (OR | 1). */
static unsigned block_bw(const uint64_t *in, uint64_t n)
{
    uint64_t o = 1;
    for (uint64_t i = 0; i < n; i++) o |= in[i];
    return oc_effective_bitwidth(o);
}

static uint64_t dvbp_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out)
{
    const uint64_t B = f->block, P = f->page, E = f->step;
    const uint64_t full_pages = round_down(n, B * P) / (B * P);
    uint64_t pos = 0, i = 0;
    for (uint64_t p = 0; p < full_pages; p++) {
        uint8_t *meta = out + pos;
        pos += P;
        for (uint64_t q = 0; q < P; q++, i += B) {
            const unsigned bw = block_bw(in + i, B);
            meta[q] = (uint8_t)bw;
            vbp_pack(in + i, B, bw, (unsigned)E, (uint64_t *)(out + pos));
            pos += B * bw / 8;
        }
    }
    if (i < n) {                                  /* dynamic_vbp.h:226-256 */
        uint8_t *meta = out + pos;
        const uint64_t nb = (n - i) / B;
        pos += P;
        for (uint64_t q = 0; q < nb; q++, i += B) {
            const unsigned bw = block_bw(in + i, B);
            meta[q] = (uint8_t)bw;
            vbp_pack(in + i, B, bw, (unsigned)E, (uint64_t *)(out + pos));
            pos += B * bw / 8;
        }
        memset(meta + nb, MS_VBP_BW_NOBLOCK, P - nb);
    }
    return pos;
}

static int dvbp_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t *used)
{
    const uint64_t B = f->block, P = f->page, E = f->step;
    uint64_t pos = 0, done = 0;
    while (done < n) {
        const uint8_t *meta = in + pos;
        pos += P;
        for (uint64_t q = 0; q < P; q++) {
            const unsigned bw = meta[q];
            if (bw == MS_VBP_BW_NOBLOCK) continue;
            if (bw < 1 || bw > 64) return MS_ERR_INVALID;
            vbp_unpack((const uint64_t *)(in + pos), B, bw, (unsigned)E, out + done);
            pos += B * bw / 8;
            done += B;
        }
    }
    *used = pos;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* Leaf formats as inner formats of a cascade                          */
/* ------------------------------------------------------------------ */

static uint64_t leaf_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out)
{
    switch (f->kind) {
    case MS_UNCOMPR: memcpy(out, in, 8 * n); return 8 * n;
    case MS_STATIC_VBP:
        vbp_pack(in, n, f->bw, f->step, (uint64_t *)out);
        return n * f->bw / 8;
    default:
        return dvbp_compress(f, in, n, out);
    }
}

static int leaf_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t *used)
{
    switch (f->kind) {
    case MS_UNCOMPR: memcpy(out, in, 8 * n); *used = 8 * n; return MS_OK;
    case MS_STATIC_VBP:
        vbp_unpack((const uint64_t *)in, n, f->bw, f->step, out);
        *used = n * f->bw / 8;
        return MS_OK;
    default:
        return dvbp_decompress(f, in, n, out, used);
    }
}

/* ------------------------------------------------------------------ */
/* delta_f (delta.h:106-266): d_i = x_i - x_{i-S}, x_{<0} = 0, carried  */
/* over blocks, calls and write-iterator flushes; each block is one    */
/* inner morph_batch call.                                             */
/* ------------------------------------------------------------------ */

static uint64_t delta_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out, uint64_t *last)
{
    ms_format_desc inner;
    const uint64_t L = f->block, S = f->step;
    uint64_t pos = 0;
    uint64_t *tmp = (uint64_t *)malloc(8 * L);
    inner_of(f, &inner);
    for (uint64_t b = 0; b < n / L; b++) {
        for (uint64_t i = 0; i < L; i++) {
            const uint64_t x = in[b * L + i];
            tmp[i] = x - last[i % S];
            last[i % S] = x;
        }
        pos += leaf_compress(&inner, tmp, L, out + pos);
    }
    free(tmp);
    return pos;
}

static int delta_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t *used)
{
    ms_format_desc inner;
    const uint64_t L = f->block, S = f->step;
    uint64_t pos = 0, u;
    uint64_t pre[64] = {0};
    inner_of(f, &inner);
    for (uint64_t b = 0; b < n / L; b++) {
        uint64_t *o = out + b * L;
        int rc = leaf_decompress(&inner, in + pos, L, o, &u);
        if (rc) return rc;
        pos += u;
        for (uint64_t i = 0; i < L; i++) {
            pre[i % S] += o[i];
            o[i] = pre[i % S];
        }
    }
    *used = pos;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* for_f (for.h:121-365)                                               */
/* ------------------------------------------------------------------ */

/* determine_reference_value (for.h:121-161): a lane-wise minimum over
 * the block's vectors, then the unsigned minimum of the lanes.  With more
 * than one lane the vector min is signed (_mm512_min_epi64, calc_avx512.h:
 * 69-78; the AVX2 min compares signed too); the scalar one is unsigned. */
static uint64_t for_ref(const uint64_t *in, uint64_t L, unsigned E)
{
    uint64_t r = U64MAX;
    if (E == 1) {
        for (uint64_t i = 0; i < L; i++) if (in[i] < r) r = in[i];
    } else {
        for (unsigned j = 0; j < E; j++) {
            int64_t m = (int64_t)in[j];
            for (uint64_t i = j + E; i < L; i += E) if ((int64_t)in[i] < m) m = (int64_t)in[i];
            if ((uint64_t)m < r) r = (uint64_t)m;
        }
    }
    if (r == U64MAX) r--;                                 /* for.h:155-158 */
    return r;
}

static uint64_t for_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out)
{
    ms_format_desc inner;
    const uint64_t L = f->block, P = f->page;
    const unsigned E = for_lanes(f);
    uint64_t pos = 0, done = 0;
    uint64_t *tmp = (uint64_t *)malloc(8 * L);
    inner_of(f, &inner);
    while (done < n) {
        uint64_t *meta = (uint64_t *)(out + pos);
        uint64_t nb = (n - done) / L;
        pos += 8 * P;
        if (nb > P) nb = P;
        if (nb < P) meta[nb] = U64MAX;                     /* for.h:214-220 */
        for (uint64_t q = 0; q < nb; q++, done += L) {
            const uint64_t ref = for_ref(in + done, L, E);
            meta[q] = ref;
            for (uint64_t i = 0; i < L; i++) tmp[i] = in[done + i] - ref;
            pos += leaf_compress(&inner, tmp, L, out + pos);
        }
    }
    free(tmp);
    return pos;
}

static int for_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t *used)
{
    ms_format_desc inner;
    const uint64_t L = f->block, P = f->page;
    uint64_t pos = 0, done = 0, u;
    inner_of(f, &inner);
    while (done < n) {
        const uint64_t *meta = (const uint64_t *)(in + pos);
        pos += 8 * P;
        for (uint64_t q = 0; q < P; q++) {
            const uint64_t ref = meta[q];
            int rc;
            if (ref == U64MAX) break;                      /* for.h:318-324 */
            rc = leaf_decompress(&inner, in + pos, L, out + done, &u);
            if (rc) return rc;
            pos += u;
            for (uint64_t i = 0; i < L; i++) out[done + i] += ref;
            done += L;
        }
    }
    *used = pos;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* rle_f (rle.h:95-147, 218-256): maximal runs as (value, length).     */
/* For n == 0 this emits no run; the reference reads in[0] and emits   */
/* one bogus run (rle.h:113-140).                                      */
/* ------------------------------------------------------------------ */

static uint64_t rle_compress(const uint64_t *in, uint64_t n, uint8_t *out8)
{
    uint64_t *out = (uint64_t *)out8, r = 0, start = 0;
    if (n == 0) return 0;
    for (uint64_t i = 1; i <= n; i++) {
        if (i == n || in[i] != in[start]) {
            out[2 * r] = in[start];
            out[2 * r + 1] = i - start;
            r++;
            start = i;
        }
    }
    return 16 * r;
}

static void rle_decompress(const uint8_t *in8, uint64_t words, uint64_t *out)
{
    const uint64_t *in = (const uint64_t *)in8;
    uint64_t o = 0;
    for (uint64_t i = 0; i + 1 < words; i += 2)
        for (uint64_t k = 0; k < in[i + 1]; k++) out[o++] = in[i];
}

/* ------------------------------------------------------------------ */
/* static_vbp_f<vbp_padding_l<bw, S>> (vbp_padding.h:52-205): a block  */
/* is 64 / bw codes per lane; lane j's word = OR_k v[k S + j] << k bw,  */
/* codes never cross words (unmasked on pack, masked on unpack).        */
/* ------------------------------------------------------------------ */
static uint64_t vbpp_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out8)
{
    const unsigned bw = f->bw, S = f->step, per = 64 / bw;
    const uint64_t bs = (uint64_t)per * S;
    uint64_t *out = (uint64_t *)out8;
    for (uint64_t b = 0; b < n / bs; b++)
        for (unsigned j = 0; j < S; j++) {
            uint64_t w = 0;
            for (unsigned k = 0; k < per; k++) w |= in[b * bs + (uint64_t)k * S + j] << (k * bw);
            out[b * S + j] = w;
        }
    return n / bs * S * 8;
}

static uint64_t vbpp_decompress(const ms_format_desc *f, const uint8_t *in8, uint64_t n, uint64_t *out)
{
    const unsigned bw = f->bw, S = f->step, per = 64 / bw;
    const uint64_t bs = (uint64_t)per * S, mask = bw_max(bw);
    const uint64_t *in = (const uint64_t *)in8;
    for (uint64_t b = 0; b < n / bs; b++)
        for (unsigned j = 0; j < S; j++)
            for (unsigned k = 0; k < per; k++) out[b * bs + (uint64_t)k * S + j] = (in[b * S + j] >> (k * bw)) & mask;
    return n / bs * S * 8;
}

/* ------------------------------------------------------------------ */
/* k_wise_ns_f<2> (k_wise_ns.h:43-262, the SSE implementation): per     */
/* pair a descriptor byte (zero_bytes(x1) << 3 | zero_bytes(x0), with   */
/* zero_bytes = leading zero bytes of (x | 1), math.h:115-120), then    */
/* the low 8 - zb bytes of x0 and of x1.                                */
/* ------------------------------------------------------------------ */
static unsigned zero_bytes(uint64_t x) { return (unsigned)__builtin_clzll(x | 1) / 8; }

static uint64_t kwns_compress(const uint64_t *in, uint64_t n, uint8_t *out)
{
    uint64_t pos = 0;
    for (uint64_t i = 0; i + 1 < n; i += 2) {
        const unsigned z0 = zero_bytes(in[i]), z1 = zero_bytes(in[i + 1]);
        out[pos++] = (uint8_t)((z1 << 3) | z0);
        for (unsigned b = 0; b < 8 - z0; b++) out[pos++] = (uint8_t)(in[i] >> (8 * b));
        for (unsigned b = 0; b < 8 - z1; b++) out[pos++] = (uint8_t)(in[i + 1] >> (8 * b));
    }
    return pos;
}

static uint64_t kwns_decompress(const uint8_t *in, uint64_t n, uint64_t *out)
{
    uint64_t pos = 0;
    for (uint64_t i = 0; i < n; i += 2) {
        const unsigned d = in[pos++], z0 = d & 7, z1 = (d >> 3) & 7;
        uint64_t x0 = 0, x1 = 0;
        for (unsigned b = 0; b < 8 - z0; b++) x0 |= (uint64_t)in[pos++] << (8 * b);
        for (unsigned b = 0; b < 8 - z1; b++) x1 |= (uint64_t)in[pos++] << (8 * b);
        out[i] = x0;
        out[i + 1] = x1;
    }
    return pos;
}

/* ------------------------------------------------------------------ */
/* group_simple_f<E, uint64_t, 8 E> (group_simple.h:64-365; packing of  */
/* group_simple_routine_gen.py:101-168).  Groups of E values, a group's */
/* pseudo maximum is their OR.  Greedily, the first mode whose next     */
/* min(n, groups left, 64) groups all fit its width takes n groups; a   */
/* mode that does not get n groups ends the selectors and the groups    */
/* left go to one incomplete block of width 64 / count.  A call writes  */
/* the selector count (uint64), the left groups (uint8), the selectors  */
/* (4 bit, low nibble first), padding to 8 E bytes from the image base  */
/* and one 8 E-byte vector per block.                                   */
/* ------------------------------------------------------------------ */
static const unsigned gs_num[15] = {64, 32, 21, 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1};
static const unsigned gs_bits[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 21, 32, 64};

static uint64_t align_to(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

/* base = the call's offset from the image base (the padding aligns to it) */
static uint64_t gs_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out, uint64_t base)
{
    const unsigned E = f->step;
    const uint64_t G = n / E;
    uint64_t *gmax = (uint64_t *)malloc(8 * (G + 1));
    uint8_t *sel = (uint8_t *)malloc(G + 1);
    uint64_t p = 0, cs = 0, data, w;
    for (uint64_t g = 0; g < G; g++) {
        uint64_t o = 0;
        for (unsigned j = 0; j < E; j++) o |= in[g * E + j];
        gmax[g] = o;
    }
    while (p < G) {
        const uint64_t avail = G - p < 64 ? G - p : 64;
        unsigned i, off = 0, nn = 0;
        for (i = 0; i < 15; i++) {
            const uint64_t mask = bw_max(gs_bits[i]);
            const uint64_t maxpos = gs_num[i] < avail ? gs_num[i] : avail;
            nn = gs_num[i];
            off = 0;
            while (off < maxpos && gmax[p + off] <= mask) off++;
            if (off == maxpos) break;
        }
        if (off != nn) break;
        sel[cs++] = (uint8_t)i;
        p += nn;
    }
    *(uint64_t *)out = cs;
    out[8] = (uint8_t)(G - p);
    for (uint64_t k = 0; k < cs; k++) {
        if (k % 2 == 0) out[9 + k / 2] = sel[k];
        else out[9 + k / 2] |= (uint8_t)(sel[k] << 4);
    }
    data = align_to(base + 9 + (cs + 1) / 2, 8ull * E) - base;
    w = data;
    p = 0;
    for (uint64_t k = 0; k < cs; k++) {
        const unsigned nn = gs_num[sel[k]], b = gs_bits[sel[k]];
        for (unsigned j = 0; j < E; j++) {
            uint64_t v = 0;
            for (unsigned q = 0; q < nn; q++) v |= in[(p + q) * E + j] << (q * b);
            memcpy(out + w + 8 * j, &v, 8);
        }
        p += nn;
        w += 8ull * E;
    }
    if (G - p) {
        const unsigned R = (unsigned)(G - p), bw = 64 / R;
        for (unsigned j = 0; j < E; j++) {
            uint64_t v = in[p * E + j];
            for (unsigned q = 1; q < R; q++) v |= in[(p + q) * E + j] << (q * bw);
            memcpy(out + w + 8 * j, &v, 8);
        }
        w += 8ull * E;
    }
    free(gmax);
    free(sel);
    return w;
}

static int gs_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t base,
                         uint64_t *used)
{
    const unsigned E = f->step;
    uint64_t done = 0, pos = 0;
    while (done < n) {
        uint64_t cs;
        memcpy(&cs, in + pos, 8);
        const unsigned R = in[pos + 8];
        const uint8_t *sel = in + pos + 9;
        uint64_t w = align_to(base + pos + 9 + (cs + 1) / 2, 8ull * E) - base;
        if (R >= 64) return MS_ERR_INVALID;
        for (uint64_t k = 0; k < cs; k++) {
            const unsigned s = (sel[k / 2] >> (4 * (k % 2))) & 15;
            if (s > 14) return MS_ERR_INVALID;
            const unsigned nn = gs_num[s], b = gs_bits[s];
            if (done + (uint64_t)nn * E > n) return MS_ERR_INVALID;
            for (unsigned j = 0; j < E; j++) {
                uint64_t v;
                memcpy(&v, in + w + 8 * j, 8);
                for (unsigned q = 0; q < nn; q++)   /* the last group is not masked */
                    out[done + (uint64_t)q * E + j] = q + 1 < nn ? (v >> (q * b)) & bw_max(b) : v >> (q * b);
            }
            done += (uint64_t)nn * E;
            w += 8ull * E;
        }
        if (R) {
            const unsigned bw = 64 / R;
            if (done + (uint64_t)R * E > n) return MS_ERR_INVALID;
            for (unsigned j = 0; j < E; j++) {
                uint64_t v;
                memcpy(&v, in + w + 8 * j, 8);
                for (unsigned q = 0; q < R; q++) out[done + (uint64_t)q * E + j] = (v >> (q * bw)) & bw_max(bw);
            }
            done += (uint64_t)R * E;
            w += 8ull * E;
        }
        if (!cs && !R) return MS_ERR_INVALID;   /* an empty call: no progress */
        pos = w;
    }
    *used = pos;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* morph_batch                                                          */
/* ------------------------------------------------------------------ */

static int morph_batch_compress_at(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out,
                                   uint64_t base, uint64_t *out_bytes);

/* ------------------------------------------------------------------ */
/* blockwise_individual_f<B, selector, ve> (blockwise_individual.h:     */
/* 113-600) with the format codes of format_selector_commons.h:25-103.  */
/* Per block of B values: the code byte, zero padding to the next       */
/* 64-byte boundary of the image (create_aligned_ptr with the 64-byte   */
/* vector size of the 8-lane processing style), then one morph_batch    */
/* call of the code's default format (default_formats.h:35-94 with 8    */
/* lanes) over the B values.  Codes 68 and 70 name macros that          */
/* default_formats.h does not define; they are refused.  Code 71 is     */
/* this library's own extension: rle_f's (value, length) pairs of the   */
/* block, which delimit themselves by their lengths.                    */
/* ------------------------------------------------------------------ */

int oc_bwi_inner(unsigned code, ms_format_desc *in)
{
    memset(in, 0, sizeof *in);
    if (code == 0) { in->kind = MS_UNCOMPR; return MS_OK; }
    if (code <= 64) { in->kind = MS_STATIC_VBP; in->bw = code; in->step = 8; return MS_OK; }
    switch (code) {
    case 65: in->kind = MS_DYNAMIC_VBP; in->block = 512; in->page = 64; in->step = 8; return MS_OK;
    case 66: in->kind = MS_GROUP_SIMPLE; in->step = 8; return MS_OK;
    case 67:
    case 69:
        in->kind = code == 67 ? MS_DELTA : MS_FOR;
        in->block = 1024;
        in->page = code == 67 ? 0 : 8;
        in->step = 8;
        in->inner_kind = MS_DYNAMIC_VBP;
        in->inner_block = 512;
        in->inner_page = 64;
        in->inner_step = 8;
        return MS_OK;
    case 71: in->kind = MS_RLE; return MS_OK;
    default: return MS_ERR_INVALID;
    }
}

/* blockwise_individual_helper::get_size_max_byte_single_block
 * (blockwise_individual.h:57-79) over the defined default formats. */
static uint64_t bwi_single_max(uint64_t B)
{
    static const unsigned codes[] = {0, 64, 65, 66, 67, 69};
    uint64_t m = 0, x;
    ms_format_desc in;
    for (unsigned i = 0; i < sizeof codes / sizeof codes[0]; i++) {
        oc_bwi_inner(codes[i], &in);
        if (oc_size_max_byte(&in, B, &x) == MS_OK && x > m) m = x;
    }
    return m;
}

/* exhaustive_size_format_selector.h:80-141 (best) and its worst-size
 * variant; the static selector of dummy_format_selectors.h:20-39. */
static int bwi_choose(const ms_format_desc *f, const uint64_t *in, uint64_t B, uint8_t *tmp, unsigned *code)
{
    static const unsigned cand[] = {65, 66, 67, 69};
    const int best = f->bw == MS_BWI_SEL_BEST;
    uint64_t size, x;
    unsigned c, i;
    ms_format_desc d;
    if (f->bw == MS_BWI_SEL_STATIC) { *code = f->page; return MS_OK; }
    {
        const unsigned mbw = block_bw(in, B);
        if (best && mbw < 64) { size = (uint64_t)mbw * B / 8; c = mbw; }
        else { size = 8 * B; c = 0; }
    }
    for (i = 0; i < sizeof cand / sizeof cand[0]; i++) {
        oc_bwi_inner(cand[i], &d);
        oc_size_max_byte(&d, B, &x);
        memset(tmp, 0, x + 64);
        if (morph_batch_compress_at(&d, in, B, tmp, 0, &x) != MS_OK) return MS_ERR_INVALID;
        if (best == (x < size)) { size = x; c = cand[i]; }
    }
    *code = c;
    return MS_OK;
}

/* `codes`: one per block, or NULL to run the format's selector; `chosen`
 * (may be NULL) receives the codes used. */
static int bwi_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, const uint8_t *codes, uint8_t *out,
                        uint64_t base, uint64_t *out_bytes, uint8_t *chosen)
{
    const uint64_t B = f->block, nb = n / B;
    uint64_t pos = 0, x;
    uint8_t *tmp = NULL;
    ms_format_desc d;
    if (!codes) {
        tmp = (uint8_t *)malloc(bwi_single_max(B) + 128);
        if (!tmp) return MS_ERR_NOMEM;
    }
    for (uint64_t b = 0; b < nb; b++) {
        unsigned code;
        int rc = MS_OK;
        if (codes) code = codes[b];
        else rc = bwi_choose(f, in + b * B, B, tmp, &code);
        if (rc == MS_OK) rc = oc_bwi_inner(code, &d);
        if (rc != MS_OK) { free(tmp); return rc; }
        if (chosen) chosen[b] = (uint8_t)code;
        out[pos] = (uint8_t)code;
        pos = align64(base + pos + 1) - base;
        if (d.kind == MS_RLE) x = rle_compress(in + b * B, B, out + pos);
        else if ((rc = morph_batch_compress_at(&d, in + b * B, B, out + pos, base + pos, &x)) != MS_OK) {
            free(tmp);
            return rc;
        }
        pos += x;
    }
    free(tmp);
    *out_bytes = pos;
    return MS_OK;
}

static int bwi_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t base,
                          uint64_t *used)
{
    const uint64_t B = f->block;
    uint64_t pos = 0, x;
    ms_format_desc d;
    for (uint64_t b = 0; b < n / B; b++) {
        const unsigned code = in[pos];
        int rc;
        if ((rc = oc_bwi_inner(code, &d)) != MS_OK) return rc;
        pos = align64(base + pos + 1) - base;
        if (d.kind == MS_RLE) {
            /* pairs until the lengths cover the block */
            uint64_t done = 0;
            x = 0;
            while (done < B) {
                uint64_t v, len;
                memcpy(&v, in + pos + x, 8);
                memcpy(&len, in + pos + x + 8, 8);
                if (len == 0 || len > B - done) return MS_ERR_INVALID;
                for (uint64_t i = 0; i < len; i++) out[b * B + done + i] = v;
                done += len;
                x += 16;
            }
        } else if (d.kind == MS_GROUP_SIMPLE) {
            if ((rc = gs_decompress(&d, in + pos, B, out + b * B, base + pos, &x)) != MS_OK) return rc;
        } else if ((rc = oc_morph_batch_decompress(&d, in + pos, B, out + b * B, &x)) != MS_OK) {
            return rc;
        }
        pos += x;
    }
    *used = pos;
    return MS_OK;
}

int oc_morph_batch_compress(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out, uint64_t *out_bytes)
{
    return morph_batch_compress_at(f, in, n, out, 0, out_bytes);
}

/* base: the call's offset from the image base (group_simple aligns its
 * data area to it) */
static int morph_batch_compress_at(const ms_format_desc *f, const uint64_t *in, uint64_t n, uint8_t *out,
                                   uint64_t base, uint64_t *out_bytes)
{
    uint64_t bs, last[64] = {0};
    if (!valid(f)) return MS_ERR_INVALID;
    if (f->kind != MS_RLE) {
        oc_block_size(f, &bs);
        if (n % bs) return MS_ERR_INVALID;
    }
    switch (f->kind) {
    case MS_UNCOMPR:
    case MS_STATIC_VBP:
    case MS_DYNAMIC_VBP: *out_bytes = leaf_compress(f, in, n, out); return MS_OK;
    case MS_DELTA: *out_bytes = delta_compress(f, in, n, out, last); return MS_OK;
    case MS_FOR: *out_bytes = for_compress(f, in, n, out); return MS_OK;
    case MS_RLE: *out_bytes = rle_compress(in, n, out); return MS_OK;
    case MS_STATIC_VBP_PAD: *out_bytes = vbpp_compress(f, in, n, out); return MS_OK;
    case MS_K_WISE_NS: *out_bytes = kwns_compress(in, n, out); return MS_OK;
    case MS_GROUP_SIMPLE: *out_bytes = gs_compress(f, in, n, out, base); return MS_OK;
    case MS_BLOCKWISE_INDIVIDUAL: return bwi_compress(f, in, n, NULL, out, base, out_bytes, NULL);
    default: return MS_ERR_INVALID;
    }
}

int oc_morph_batch_decompress(const ms_format_desc *f, const uint8_t *in, uint64_t n, uint64_t *out, uint64_t *in_bytes)
{
    uint64_t bs;
    if (!valid(f)) return MS_ERR_INVALID;
    if (f->kind != MS_RLE && (oc_block_size(f, &bs) || n % bs)) return MS_ERR_INVALID;
    switch (f->kind) {
    case MS_STATIC_VBP_PAD: *in_bytes = vbpp_decompress(f, in, n, out); return MS_OK;
    case MS_K_WISE_NS: *in_bytes = kwns_decompress(in, n, out); return MS_OK;
    case MS_GROUP_SIMPLE: return gs_decompress(f, in, n, out, 0, in_bytes);
    case MS_BLOCKWISE_INDIVIDUAL: return bwi_decompress(f, in, n, out, 0, in_bytes);
    case MS_UNCOMPR:
    case MS_STATIC_VBP:
    case MS_DYNAMIC_VBP: return leaf_decompress(f, in, n, out, in_bytes);
    case MS_DELTA: return delta_decompress(f, in, n, out, in_bytes);
    case MS_FOR: return for_decompress(f, in, n, out, in_bytes);
    default: return MS_ERR_INVALID;
    }
}

/* ------------------------------------------------------------------ */
/* Columns                                                             */
/* ------------------------------------------------------------------ */

static int col_alloc(oc_col *c, const ms_format_desc *f, uint64_t size_alloc)
{
    memset(c, 0, sizeof *c);
    c->f = *f;
    c->size_alloc = size_alloc;
    c->data = (uint8_t *)calloc(size_alloc + 64, 1);
    return c->data ? MS_OK : MS_ERR_NOMEM;
}

void oc_col_free(oc_col *c)
{
    if (c && c->data) free(c->data);
    if (c) c->data = NULL;
}

int oc_col_from_values(oc_col *out, const uint64_t *v, uint64_t n)
{
    ms_format_desc u;
    memset(&u, 0, sizeof u);
    u.kind = MS_UNCOMPR;
    if (col_alloc(out, &u, 8 * n)) return MS_ERR_NOMEM;
    if (n) memcpy(out->data, v, 8 * n);
    out->count = n;
    out->size_used = 8 * n;          /* set_meta_data(count, size): sizeCompr = 0 */
    out->size_compr = 0;
    return MS_OK;
}

/* A column over a copy of an existing byte image (for timing the CPU
 * baseline on the exact bytes the GPU produced). */
int oc_col_from_image(oc_col *out, const ms_format_desc *f, const uint8_t *image, uint64_t count,
                      uint64_t size_used, uint64_t size_compr)
{
    if (!valid(f)) return MS_ERR_INVALID;
    if (col_alloc(out, f, size_used + 64)) return MS_ERR_NOMEM;
    if (size_used) memcpy(out->data, image, size_used);
    out->count = count;
    out->size_used = size_used;
    out->size_compr = size_compr;
    return MS_OK;
}

/* column.h:205-214 */
uint64_t oc_count_values_uncompr(const oc_col *c)
{
    uint64_t r;
    if (c->size_compr == c->size_used) return 0;
    if (oc_rest_offset(&c->f, c->size_compr, &r)) return 0;
    return (c->size_used - r) / 8;
}

/* morph.h:157-197 (compression), rle.h:155-188 */
int oc_morph_compress(oc_col *out, const ms_format_desc *dst, const uint64_t *in, uint64_t n)
{
    uint64_t alloc, bs, nc, c = 0, r;
    if (!valid(dst)) return MS_ERR_INVALID;
    if (dst->kind == MS_UNCOMPR) return oc_col_from_values(out, in, n);
    if (dst->kind == MS_BLOCKWISE_INDIVIDUAL) return oc_morph_compress_codes(out, dst, in, n, NULL);
    if (dst->kind == MS_RLE) {
        if (col_alloc(out, dst, 16 * n + 16)) return MS_ERR_NOMEM;
        c = rle_compress(in, n, out->data);
        out->count = n;
        out->size_used = out->size_compr = c;
        return MS_OK;
    }
    oc_size_max_byte_any_len(dst, n, &alloc);
    if (col_alloc(out, dst, alloc)) return MS_ERR_NOMEM;
    oc_block_size(dst, &bs);
    nc = round_down(n, bs);
    if (nc) oc_morph_batch_compress(dst, in, nc, out->data, &c);
    out->count = n;
    out->size_compr = c;
    if (n - nc) {
        oc_rest_offset(dst, c, &r);
        memcpy(out->data + r, in + nc, 8 * (n - nc));
        out->size_used = r + 8 * (n - nc);
    } else {
        out->size_used = c;
    }
    return MS_OK;
}

/* morph<ve, blockwise_individual_f, uncompr_f> at column level with the
 * block codes given (one per block) or, codes == NULL, chosen by the
 * format's selector. */
int oc_morph_compress_codes(oc_col *out, const ms_format_desc *dst, const uint64_t *in, uint64_t n,
                            const uint8_t *codes)
{
    uint64_t alloc, nc, c = 0, r;
    int rc;
    if (!valid(dst) || dst->kind != MS_BLOCKWISE_INDIVIDUAL) return MS_ERR_INVALID;
    oc_size_max_byte_any_len(dst, n, &alloc);
    alloc += n / dst->block * 8 * dst->block;   /* rle_f views (code 71) take up to 16 bytes per value */
    if (col_alloc(out, dst, alloc)) return MS_ERR_NOMEM;
    nc = round_down(n, dst->block);
    if (nc && (rc = bwi_compress(dst, in, nc, codes, out->data, 0, &c, NULL)) != MS_OK) {
        oc_col_free(out);
        return rc;
    }
    out->count = n;
    out->size_compr = c;
    if (n - nc) {
        oc_rest_offset(dst, c, &r);
        memcpy(out->data + r, in + nc, 8 * (n - nc));
        out->size_used = r + 8 * (n - nc);
    } else {
        out->size_used = c;
    }
    return MS_OK;
}

/* The code byte of every block of a blockwise_individual_f column. */
int oc_bwi_codes(const oc_col *c, uint8_t *codes, uint64_t cap, uint64_t *nblocks)
{
    uint64_t pos = 0, x;
    const uint64_t B = c->f.block, nb = (c->count - oc_count_values_uncompr(c)) / B;
    ms_format_desc d;
    if (c->f.kind != MS_BLOCKWISE_INDIVIDUAL) return MS_ERR_INVALID;
    *nblocks = nb;
    for (uint64_t b = 0; b < nb; b++) {
        uint64_t *tmp;
        int rc;
        if (pos >= c->size_compr) return MS_ERR_INVALID;
        if (b < cap) codes[b] = c->data[pos];
        if ((rc = oc_bwi_inner(c->data[pos], &d)) != MS_OK) return rc;
        /* the view's size: decompress it */
        tmp = (uint64_t *)malloc(8 * B);
        if (!tmp) return MS_ERR_NOMEM;
        {
            ms_format_desc one = c->f;
            one.bw = MS_BWI_SEL_STATIC;
            rc = bwi_decompress(&one, c->data + pos, B, tmp, pos, &x);
        }
        free(tmp);
        if (rc != MS_OK) return rc;
        pos += x;
    }
    return MS_OK;
}

/* The exhaustive (or static) selector of a blockwise_individual_f format
 * over one block of B values. */
int oc_bwi_choose(const ms_format_desc *f, const uint64_t *in, unsigned *code)
{
    uint8_t *tmp;
    int rc;
    if (!valid(f) || f->kind != MS_BLOCKWISE_INDIVIDUAL) return MS_ERR_INVALID;
    tmp = (uint8_t *)malloc(bwi_single_max(f->block) + 128);
    if (!tmp) return MS_ERR_NOMEM;
    rc = bwi_choose(f, in, f->block, tmp, code);
    free(tmp);
    return rc;
}

/* morph.h:203-241 (decompression), rle.h:264-297 */
int oc_morph_decompress(const oc_col *in, uint64_t *out)
{
    uint64_t rest, nc, used, r;
    if (!valid(&in->f)) return MS_ERR_INVALID;
    if (in->f.kind == MS_UNCOMPR) { memcpy(out, in->data, 8 * in->count); return MS_OK; }
    if (in->f.kind == MS_RLE) { rle_decompress(in->data, in->size_used / 8, out); return MS_OK; }
    rest = oc_count_values_uncompr(in);
    nc = in->count - rest;
    if (nc) {
        int rc = oc_morph_batch_decompress(&in->f, in->data, nc, out, &used);
        if (rc) return rc;
    }
    if (rest) {
        oc_rest_offset(&in->f, in->size_compr, &r);
        memcpy(out + nc, in->data + r, 8 * rest);
    }
    return MS_OK;
}

/* column.h:241-268: the uncompressed rest, zero-padded to a multiple of
 * the block size, is compressed by one morph_batch call (fresh state: a
 * new dynamic_vbp page, for_f reference page, zero delta carry) to the end
 * of the compressed part, inside the gap; the meta data stay as they are.
 * No action for uncompr_f or without a rest (rle_f has no block size and
 * is refused). */
int oc_prepare_for_random_access(oc_col *c, int *acted)
{
    uint64_t rest, bs, ru, r, b;
    *acted = 0;
    if (c->f.kind == MS_RLE) return MS_ERR_INVALID;
    if (c->f.kind == MS_UNCOMPR) return MS_OK;
    rest = oc_count_values_uncompr(c);
    if (!rest) return MS_OK;
    oc_block_size(&c->f, &bs);
    ru = round_up(rest, bs);
    oc_rest_offset(&c->f, c->size_compr, &r);
    /* column.h:251-262 writes without a check; the buffer holds size_alloc
     * plus the 64 bytes of alignment slack a column allocates */
    if (r + 8 * ru > c->size_alloc + 64) return MS_ERR_CAPACITY;
    memset(c->data + r + 8 * rest, 0, 8 * (ru - rest));
    memset(c->data + c->size_compr, 0, r - c->size_compr);       /* packing ORs into zeroed words */
    morph_batch_compress_at(&c->f, (const uint64_t *)(c->data + r), ru, c->data + c->size_compr, c->size_compr, &b);
    if (c->size_compr + b > r) return MS_ERR_CAPACITY;
    *acted = 1;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* Write-iterator output (write_iterator.h:66-196; delta.h:369-490 for */
/* the delta carry; uncompr.h:125-221 for the unbuffered uncompr one)  */
/* ------------------------------------------------------------------ */

int oc_wit_emit(oc_col *out, const ms_format_desc *f, const uint64_t *v, uint64_t m,
                uint64_t extra_capacity, uint64_t *append_off)
{
    uint64_t alloc, bs, K, pos = 0, full, rem, nc, last[64] = {0}, r, b;
    if (!valid(f) || f->kind == MS_RLE) return MS_ERR_INVALID;   /* rle_f has no write-iterator */
    if (f->kind == MS_UNCOMPR) {
        if (col_alloc(out, f, 8 * m + extra_capacity)) return MS_ERR_NOMEM;
        if (m) memcpy(out->data, v, 8 * m);
        out->count = m;
        out->size_used = 8 * m;
        out->size_compr = 0;
        *append_off = 8 * m;
        return MS_OK;
    }
    oc_block_size(f, &bs);
    K = round_up(bs, 2048);                                      /* m_CountBuffer */
    /* every K-value morph_batch call starts fresh pages, so the calls are
     * bounded one by one (at most m / K full calls and one shorter one,
     * get_size_max_byte being monotone in its count), then the raw rest */
    oc_size_max_byte(f, K, &alloc);
    alloc = (m / K + 1) * alloc + 8 * bs;
    oc_rest_offset(f, 0, &r);                  /* the rest sits behind the gap */
    if (col_alloc(out, f, alloc + r + extra_capacity + 2 * 64)) return MS_ERR_NOMEM;
    full = m / K;
    for (uint64_t c = 0; c < full; c++) {
        if (f->kind == MS_DELTA) pos += delta_compress(f, v + c * K, K, out->data + pos, last);
        else { oc_morph_batch_compress(f, v + c * K, K, out->data + pos, &b); pos += b; }
    }
    rem = m - full * K;
    out->count = m;
    if (rem) {
        nc = round_down(rem, bs);
        if (nc) {
            if (f->kind == MS_DELTA) pos += delta_compress(f, v + full * K, nc, out->data + pos, last);
            else { oc_morph_batch_compress(f, v + full * K, nc, out->data + pos, &b); pos += b; }
        }
        out->size_compr = pos;
        oc_rest_offset(f, pos, &r);
        if (rem - nc) {
            memcpy(out->data + r, v + full * K + nc, 8 * (rem - nc));
            out->size_used = r + 8 * (rem - nc);
            *append_off = out->size_used;
        } else {
            out->size_used = pos;
            *append_off = r;
        }
    } else {
        out->size_compr = pos;
        out->size_used = pos;
        oc_rest_offset(f, pos, append_off);
    }
    if (out->size_used > out->size_alloc) return MS_ERR_CAPACITY;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* Operators                                                           */
/* ------------------------------------------------------------------ */

static int cmp_apply(uint32_t cmp, uint64_t x, uint64_t val)
{
    switch (cmp) {
    case MS_CMP_EQ: return x == val;
    case MS_CMP_NE: return x != val;
    case MS_CMP_LT: return x < val;
    case MS_CMP_LE: return x <= val;
    case MS_CMP_GT: return x > val;
    case MS_CMP_GE: return x >= val;
    default: return 0;
    }
}

static int decode_all(const oc_col *in, uint64_t **vals)
{
    *vals = (uint64_t *)malloc(8 * (in->count + 1));
    if (!*vals) return MS_ERR_NOMEM;
    return oc_morph_decompress(in, *vals);
}

/* select_compr.h:108-259 */
int oc_select(oc_col *out, uint32_t cmp, uint64_t val, const ms_format_desc *out_f,
              const oc_col *in, uint32_t lanes)
{
    uint64_t *x, *pos, na = 0, nb = 0, app;
    uint64_t n = in->count, rest = oc_count_values_uncompr(in), nvec;
    int has_rest = in->size_compr != in->size_used, rc;
    if (cmp > MS_CMP_GE || lanes == 0 || lanes > 64) return MS_ERR_INVALID;
    rc = decode_all(in, &x);
    if (rc) { free(x); return rc; }
    pos = (uint64_t *)malloc(8 * (n + 1));
    nvec = has_rest ? (n - rest) + round_down(rest, lanes) : n;
    for (uint64_t i = 0; i < nvec; i++) if (cmp_apply(cmp, x[i], val)) pos[na++] = i;
    rc = oc_wit_emit(out, out_f, pos, na, 8 * lanes + 64, &app);
    if (!rc && has_rest && rest % lanes) {
        uint64_t *dst = (uint64_t *)(out->data + app);
        for (uint64_t i = nvec; i < n; i++) if (cmp_apply(cmp, x[i], val)) dst[nb++] = i;
        out->size_used = app + 8 * nb;
        out->count = na + nb;
    }
    free(x);
    free(pos);
    return rc;
}

/* between_compr.h:107-270: select with two comparisons, same write-
 * iterator output and scalar-remainder append as select. */
int oc_between(oc_col *out, uint32_t cmp_lo, uint64_t val_lo, uint32_t cmp_hi, uint64_t val_hi,
               const ms_format_desc *out_f, const oc_col *in, uint32_t lanes)
{
    uint64_t *x, *pos, na = 0, nb = 0, app;
    uint64_t n = in->count, rest = oc_count_values_uncompr(in), nvec;
    int has_rest = in->size_compr != in->size_used, rc;
    if (cmp_lo > MS_CMP_GE || cmp_hi > MS_CMP_GE || lanes == 0 || lanes > 64) return MS_ERR_INVALID;
    rc = decode_all(in, &x);
    if (rc) { free(x); return rc; }
    pos = (uint64_t *)malloc(8 * (n + 1));
    nvec = has_rest ? (n - rest) + round_down(rest, lanes) : n;
    for (uint64_t i = 0; i < nvec; i++)
        if (cmp_apply(cmp_lo, x[i], val_lo) && cmp_apply(cmp_hi, x[i], val_hi)) pos[na++] = i;
    rc = oc_wit_emit(out, out_f, pos, na, 8 * lanes + 64, &app);
    if (!rc && has_rest && rest % lanes) {
        uint64_t *dst = (uint64_t *)(out->data + app);
        for (uint64_t i = nvec; i < n; i++)
            if (cmp_apply(cmp_lo, x[i], val_lo) && cmp_apply(cmp_hi, x[i], val_hi)) dst[nb++] = i;
        out->size_used = app + 8 * nb;
        out->count = na + nb;
    }
    free(x);
    free(pos);
    return rc;
}

/* ------------------------------------------------------------------ */
/* group and join (scalar/group_uncompr.h:44-113,                       */
/* scalar/join_uncompr.h:44-128): the reference's scalar algorithms,    */
/* a hash map for group ids in order of first appearance and nested     */
/* loops for the joins.                                                 */
/* ------------------------------------------------------------------ */
typedef struct gmap {
    uint64_t *ka, *kb, *id;   /* keys (ka, kb) -> id + 1 (0: empty) */
    uint64_t cap;
} gmap;

static uint64_t mix64(uint64_t a, uint64_t b)
{
    uint64_t z = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull + (a << 6) + (a >> 2));
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* groupIds[key]: the slot of key, inserted with id 0 when new */
static uint64_t *gmap_slot(gmap *m, uint64_t a, uint64_t b)
{
    uint64_t h = mix64(a, b) & (m->cap - 1);
    while (m->id[h] && (m->ka[h] != a || m->kb[h] != b)) h = (h + 1) & (m->cap - 1);
    if (!m->id[h]) { m->ka[h] = a; m->kb[h] = b; }
    return &m->id[h];
}

int oc_group(oc_col *out_gr, oc_col *out_ext, const oc_col *in_gr, const oc_col *in_data)
{
    uint64_t *x, *g = NULL, n = in_data->count, next = 0, next_ext = 0;
    ms_format_desc u;
    gmap m;
    int rc;
    if (in_gr && in_gr->count != n) return MS_ERR_SIZE;   /* group_uncompr.h:52-57 */
    rc = decode_all(in_data, &x);
    if (rc) { free(x); return rc; }
    if (in_gr && (rc = decode_all(in_gr, &g))) { free(x); free(g); return rc; }
    memset(&u, 0, sizeof u);
    u.kind = MS_UNCOMPR;
    m.cap = 16;
    while (m.cap < 2 * n + 16) m.cap <<= 1;
    m.ka = (uint64_t *)calloc(m.cap, 8);
    m.kb = (uint64_t *)calloc(m.cap, 8);
    m.id = (uint64_t *)calloc(m.cap, 8);
    if (col_alloc(out_gr, &u, 8 * n + 8) || col_alloc(out_ext, &u, 8 * n + 8)) {
        rc = MS_ERR_NOMEM;
    } else {
        uint64_t *og = (uint64_t *)out_gr->data, *oe = (uint64_t *)out_ext->data;
        for (uint64_t i = 0; i < n; i++) {
            uint64_t *id = gmap_slot(&m, g ? g[i] : 0, x[i]);
            if (!*id) {                 /* the key was not found */
                *id = ++next;
                oe[next_ext++] = i;
            }
            og[i] = *id - 1;
        }
        out_gr->count = n;
        out_gr->size_used = 8 * n;
        out_ext->count = next_ext;
        out_ext->size_used = 8 * next_ext;
    }
    free(m.ka); free(m.kb); free(m.id); free(x); free(g);
    return rc;
}

/* nested_loop_join: every equal pair, the outer loop over the larger input
 * (the left one when both are equally long), positions ascending */
int oc_join(oc_col *out_l, oc_col *out_r, const oc_col *l, const oc_col *r)
{
    uint64_t *a, *b, na = l->count, nb = r->count, k = 0;
    ms_format_desc u;
    int rc, swap = na < nb;
    rc = decode_all(l, &a);
    if (rc) { free(a); return rc; }
    rc = decode_all(r, &b);
    if (rc) { free(a); free(b); return rc; }
    uint64_t total = 0;
    for (uint64_t i = 0; i < na; i++) for (uint64_t j = 0; j < nb; j++) total += a[i] == b[j];
    memset(&u, 0, sizeof u);
    u.kind = MS_UNCOMPR;
    if (col_alloc(out_l, &u, 8 * total + 8) || col_alloc(out_r, &u, 8 * total + 8)) {
        free(a); free(b);
        return MS_ERR_NOMEM;
    }
    uint64_t *ol = (uint64_t *)out_l->data, *orr = (uint64_t *)out_r->data;
    if (!swap) {
        for (uint64_t i = 0; i < na; i++)
            for (uint64_t j = 0; j < nb; j++)
                if (a[i] == b[j]) { ol[k] = i; orr[k] = j; k++; }
    } else {                               /* join(r, l) with the outputs swapped back */
        for (uint64_t j = 0; j < nb; j++)
            for (uint64_t i = 0; i < na; i++)
                if (a[i] == b[j]) { ol[k] = i; orr[k] = j; k++; }
    }
    out_l->count = out_r->count = k;
    out_l->size_used = out_r->size_used = 8 * k;
    free(a); free(b);
    return MS_OK;
}

/* left_semi_nto1_nested_loop_join: ascending left positions whose value
 * occurs in the right input */
int oc_semi_join(oc_col *out, const oc_col *l, const oc_col *r)
{
    uint64_t *a, *b, na = l->count, nb = r->count, k = 0;
    ms_format_desc u;
    int rc = decode_all(l, &a);
    if (rc) { free(a); return rc; }
    rc = decode_all(r, &b);
    if (rc) { free(a); free(b); return rc; }
    memset(&u, 0, sizeof u);
    u.kind = MS_UNCOMPR;
    if (col_alloc(out, &u, 8 * na + 8)) { free(a); free(b); return MS_ERR_NOMEM; }
    uint64_t *o = (uint64_t *)out->data;
    for (uint64_t i = 0; i < na; i++)
        for (uint64_t j = 0; j < nb; j++)
            if (a[i] == b[j]) { o[k++] = i; break; }
    out->count = k;
    out->size_used = 8 * k;
    free(a); free(b);
    return MS_OK;
}

/* project_compr.h:80-261 (random access replaced by decompression) */
int oc_project(oc_col *out, const ms_format_desc *out_f, const oc_col *data,
               const oc_col *pos, uint32_t lanes)
{
    uint64_t *x, *p, *v, app, np = pos->count, rest = oc_count_values_uncompr(pos), nvec;
    int has_rest = pos->size_compr != pos->size_used, rc;
    if (lanes == 0 || lanes > 64) return MS_ERR_INVALID;
    if ((rc = decode_all(data, &x))) { free(x); return rc; }
    if ((rc = decode_all(pos, &p))) { free(x); free(p); return rc; }
    v = (uint64_t *)malloc(8 * (np + 1));
    for (uint64_t i = 0; i < np; i++) {
        if (p[i] >= data->count) { free(x); free(p); free(v); return MS_ERR_RANGE; }
        v[i] = x[p[i]];
    }
    nvec = has_rest ? (np - rest) + round_down(rest, lanes) : np;
    rc = oc_wit_emit(out, out_f, v, nvec, 8 * lanes + 64, &app);
    if (!rc && has_rest && rest % lanes) {
        memcpy(out->data + app, v + nvec, 8 * (np - nvec));
        out->size_used = app + 8 * (np - nvec);
        out->count = np;
    }
    free(x); free(p); free(v);
    return rc;
}

/* agg_sum_compr.h:78-153, general_vectorized/agg_sum_uncompr.h:102-129 */
int oc_agg_sum(oc_col *out, const oc_col *in)
{
    uint64_t *x, s = 0;
    ms_format_desc u;
    int rc = decode_all(in, &x);
    if (rc) { free(x); return rc; }
    for (uint64_t i = 0; i < in->count; i++) s += x[i];
    free(x);
    memset(&u, 0, sizeof u);
    u.kind = MS_UNCOMPR;
    if (col_alloc(out, &u, 8)) return MS_ERR_NOMEM;
    memcpy(out->data, &s, 8);
    out->count = 1;
    out->size_used = 8;
    out->size_compr = in->f.kind == MS_UNCOMPR ? 0 : 8;
    return MS_OK;
}

/* scalar/intersect_uncompr.h:41-86 */
int oc_intersect_sorted(oc_col *out, const ms_format_desc *out_f, const oc_col *l, const oc_col *r)
{
    uint64_t *a, *b, *o, i = 0, j = 0, k = 0, app;
    int rc;
    if ((rc = decode_all(l, &a))) { free(a); return rc; }
    if ((rc = decode_all(r, &b))) { free(a); free(b); return rc; }
    o = (uint64_t *)malloc(8 * (l->count + 1));
    while (i < l->count && j < r->count) {
        if (a[i] < b[j]) i++;
        else if (b[j] < a[i]) j++;
        else { o[k++] = a[i]; i++; j++; }
    }
    rc = oc_wit_emit(out, out_f, o, k, 0, &app);
    free(a); free(b); free(o);
    return rc;
}

/* scalar/merge_uncompr.h:39-103 */
int oc_merge_sorted(oc_col *out, const ms_format_desc *out_f, const oc_col *l, const oc_col *r)
{
    uint64_t *a, *b, *o, i = 0, j = 0, k = 0, app;
    int rc;
    if ((rc = decode_all(l, &a))) { free(a); return rc; }
    if ((rc = decode_all(r, &b))) { free(a); free(b); return rc; }
    o = (uint64_t *)malloc(8 * (l->count + r->count + 1));
    while (i < l->count && j < r->count) {
        if (a[i] < b[j]) o[k++] = a[i++];
        else if (b[j] < a[i]) o[k++] = b[j++];
        else { o[k++] = a[i]; i++; j++; }
    }
    while (i < l->count) o[k++] = a[i++];
    while (j < r->count) o[k++] = b[j++];
    rc = oc_wit_emit(out, out_f, o, k, 0, &app);
    free(a); free(b); free(o);
    return rc;
}

static uint64_t calc_apply(uint32_t op, uint64_t a, uint64_t b)
{
    switch (op) {
    case MS_OP_ADD: return a + b;
    case MS_OP_SUB: return a - b;
    case MS_OP_MUL: return a * b;
    case MS_OP_DIV: return b ? a / b : 0;
    case MS_OP_MOD: return b ? a % b : 0;
    case MS_OP_MIN: return a < b ? a : b;
    case MS_OP_MAX: return a > b ? a : b;
    case MS_OP_AND: return a & b;
    case MS_OP_OR: return a | b;
    default: return a ^ b;
    }
}

/* scalar/calc_uncompr.h:70-106 */
int oc_calc_binary(oc_col *out, uint32_t op, const ms_format_desc *out_f, const oc_col *l, const oc_col *r)
{
    uint64_t *a, *b, app;
    int rc;
    if (op > MS_OP_XOR) return MS_ERR_INVALID;
    if (l->count != r->count) return MS_ERR_SIZE;
    if ((rc = decode_all(l, &a))) { free(a); return rc; }
    if ((rc = decode_all(r, &b))) { free(a); free(b); return rc; }
    for (uint64_t i = 0; i < l->count; i++) a[i] = calc_apply(op, a[i], b[i]);
    rc = oc_wit_emit(out, out_f, a, l->count, 0, &app);
    free(a); free(b);
    return rc;
}

int oc_select_sum(uint64_t *result, uint32_t cmp, uint64_t val, const oc_col *in)
{
    uint64_t *x, s = 0;
    int rc = decode_all(in, &x);
    if (rc) { free(x); return rc; }
    for (uint64_t i = 0; i < in->count; i++) if (cmp_apply(cmp, x[i], val)) s += x[i];
    free(x);
    *result = s;
    return MS_OK;
}

/* ------------------------------------------------------------------ */
/* Streaming fused select + sum: decodes one block at a time into a    */
/* stack buffer and folds it, like decompress_and_process_batch with   */
/* the select core feeding the agg_sum core.                           */
/* ------------------------------------------------------------------ */

typedef struct {
    const oc_col *in;
    const uint64_t *blk_off;   /* dynamic_vbp: payload offset per block  */
    const uint8_t *blk_bw;
    uint64_t b0, b1;           /* block range                             */
    uint32_t cmp;
    uint64_t val;
    uint64_t sum;
} ss_job;

static uint64_t fold(uint32_t cmp, uint64_t val, const uint64_t *x, uint64_t n)
{
    uint64_t s = 0;
    switch (cmp) {
    case MS_CMP_LT: for (uint64_t i = 0; i < n; i++) s += x[i] < val ? x[i] : 0; break;
    case MS_CMP_EQ: for (uint64_t i = 0; i < n; i++) s += x[i] == val ? x[i] : 0; break;
    default: for (uint64_t i = 0; i < n; i++) s += cmp_apply(cmp, x[i], val) ? x[i] : 0; break;
    }
    return s;
}

static void *ss_worker(void *arg)
{
    ss_job *jb = (ss_job *)arg;
    const oc_col *in = jb->in;
    uint64_t buf[4096];
    uint64_t s = 0;
    if (in->f.kind == MS_DYNAMIC_VBP) {
        const uint64_t B = in->f.block;
        for (uint64_t b = jb->b0; b < jb->b1; b++) {
            vbp_unpack((const uint64_t *)(in->data + jb->blk_off[b]), B, jb->blk_bw[b], in->f.step, buf);
            s += fold(jb->cmp, jb->val, buf, B);
        }
    } else if (in->f.kind == MS_STATIC_VBP) {
        const uint64_t B = 64ull * in->f.step, bytes = B * in->f.bw / 8;
        for (uint64_t b = jb->b0; b < jb->b1; b++) {
            vbp_unpack((const uint64_t *)(in->data + b * bytes), B, in->f.bw, in->f.step, buf);
            s += fold(jb->cmp, jb->val, buf, B);
        }
    } else {
        const uint64_t *x = (const uint64_t *)in->data;
        s = fold(jb->cmp, jb->val, x + jb->b0, jb->b1 - jb->b0);
    }
    jb->sum = s;
    return NULL;
}

int oc_select_sum_stream(uint64_t *result, uint32_t cmp, uint64_t val, const oc_col *in, int threads)
{
    uint64_t nblocks = 0, s = 0, rest, nc, r;
    uint64_t *off = NULL;
    uint8_t *bw = NULL;
    ss_job jobs[256];
    pthread_t th[256];
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (in->f.kind != MS_DYNAMIC_VBP && in->f.kind != MS_STATIC_VBP && in->f.kind != MS_UNCOMPR)
        return oc_select_sum(result, cmp, val, in);
    if ((in->f.kind == MS_DYNAMIC_VBP && in->f.block > 4096) ||
        (in->f.kind == MS_STATIC_VBP && 64ull * in->f.step > 4096))
        return oc_select_sum(result, cmp, val, in);
    rest = oc_count_values_uncompr(in);
    nc = in->count - rest;
    if (in->f.kind == MS_DYNAMIC_VBP) {
        /* walk the page headers once to find every block's payload */
        const uint64_t B = in->f.block, P = in->f.page;
        uint64_t pos = 0, done = 0;
        nblocks = nc / B;
        off = (uint64_t *)malloc(8 * (nblocks + 1));
        bw = (uint8_t *)malloc(nblocks + 1);
        nblocks = 0;
        while (done < nc) {
            const uint8_t *meta = in->data + pos;
            pos += P;
            for (uint64_t q = 0; q < P; q++) {
                if (meta[q] == MS_VBP_BW_NOBLOCK) continue;
                off[nblocks] = pos;
                bw[nblocks++] = meta[q];
                pos += B * meta[q] / 8;
                done += B;
            }
        }
    } else if (in->f.kind == MS_STATIC_VBP) {
        nblocks = nc / (64ull * in->f.step);
    } else {
        nblocks = in->count;   /* uncompr: the whole column is the rest */
        rest = 0;
    }
    for (int t = 0; t < threads; t++) {
        jobs[t].in = in;
        jobs[t].blk_off = off;
        jobs[t].blk_bw = bw;
        jobs[t].b0 = nblocks * t / threads;
        jobs[t].b1 = nblocks * (t + 1) / threads;
        jobs[t].cmp = cmp;
        jobs[t].val = val;
        if (threads == 1) ss_worker(&jobs[0]);
        else pthread_create(&th[t], NULL, ss_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) {
        if (threads > 1) pthread_join(th[t], NULL);
        s += jobs[t].sum;
    }
    if (rest) {
        oc_rest_offset(&in->f, in->size_compr, &r);
        s += fold(cmp, val, (const uint64_t *)(in->data + r), rest);
    }
    free(off);
    free(bw);
    *result = s;
    return MS_OK;
}
