// ms_encode.hip -- write side of the compression-aware path on gfx950.
//
// One emitter produces every compressed format the operators write:
//   * column-level morph (morph.h:157-197): one morph_batch call over the
//     compressed part, then the uncompressed rest behind the gap;
//   * write-iterator output (write_iterator.h:66-196, delta.h:369-490):
//     independent morph_batch calls of m_CountBuffer = round_up(bs, 2048)
//     values, each starting fresh pages, then done()'s remainder and rest.
// The innermost vertical-bit-packed blocks are the unit of work: a kernel
// finds each block's bit width (dynamic_vbp.h:122-158) and the bytes of
// page headers in front of it (dvbp meta bytes, for_f reference pages), a
// scan places the blocks, and a pack kernel assembles each block's
// interleaved words in LDS and writes payload, headers and the block index.
#include "ms_kernels.h"

namespace ms {

static constexpr int WG = 256;

static inline uint32_t grid_for(uint64_t items, uint64_t per_wg, uint32_t cap = 8192)
{
    uint64_t g = (items + per_wg - 1) / per_wg;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

struct emit_layout {
    uint32_t kind;       // MS_STATIC_VBP, MS_DYNAMIC_VBP, MS_DELTA, MS_FOR
    uint32_t leaf;       // innermost layout
    uint32_t bw;         // static leaf bit width (64 for an uncompressed leaf)
    uint32_t S;          // innermost lanes
    uint32_t dstep;      // delta step
    uint32_t E;          // for_f lanes (reference-value minimum)
    uint64_t B;          // innermost block (values)
    uint64_t P;          // dvbp leaf page (blocks)
    uint64_t L;          // outer block (delta/for) or B
    uint64_t FP;         // for_f page (blocks)
    uint64_t bpo;        // innermost blocks per outer block
    uint64_t K;          // values per morph_batch call (0: one call)
    uint64_t bs;         // outer m_BlockSize
    uint64_t gap;        // rest gap
    uint64_t cap;        // bytes of the output image
    uint64_t *status;    // the output's status word (MS_ERR_CAPACITY past cap)
    // Chunked mode (both null otherwise): every K values are one morph_batch
    // call over the values v[cmap[c] * K ...] (its delta predecessors start
    // at zero), written at cout[c] of the image.
    const uint32_t *cmap;
    const uint64_t *cout;
};

struct emit_plan {
    uint64_t m;          // values
    uint64_t m_c;        // values in the compressed part
    uint64_t nb;         // innermost blocks
    uint64_t nfb;        // outer blocks (for_f)
};

// a / d and a % d (d > 0) for the layout's block and page counts, which are
// mostly powers of two: a shift and a mask then, the division otherwise.
__device__ __forceinline__ uint64_t ldiv(uint64_t a, uint64_t d)
{
    return (d & (d - 1)) == 0 ? a >> __builtin_ctzll(d) : a / d;
}
__device__ __forceinline__ uint64_t lmod(uint64_t a, uint64_t d)
{
    return (d & (d - 1)) == 0 ? a & (d - 1) : a % d;
}

__device__ __forceinline__ emit_plan plan_of(const emit_layout &l, uint64_t m)
{
    emit_plan p;
    p.m = m;
    uint64_t mc;
    if (l.K == 0) mc = ldiv(m, l.bs) * l.bs;
    else {
        const uint64_t full = ldiv(m, l.K), rem = m - full * l.K;
        mc = full * l.K + ldiv(rem, l.bs) * l.bs;
    }
    p.m_c = mc;
    p.nb = ldiv(mc, l.B);
    p.nfb = ldiv(mc, l.L);
    return p;
}

// Records that a write would pass the end of the output image; the write
// is dropped (the host sizes outputs so that this does not happen).
__device__ __forceinline__ void emit_overflow(const emit_layout &l)
{
    atomicCAS((unsigned long long *)l.status, 0ull, (unsigned long long)MS_ERR_CAPACITY);
}

// Bytes of headers in front of innermost block b.
__device__ __forceinline__ uint64_t hdr_of(const emit_layout &l, uint64_t b)
{
    uint64_t h = 0;
    if (l.leaf == MS_DYNAMIC_VBP) {
        uint64_t q;
        if (l.kind == MS_DYNAMIC_VBP) q = l.K ? lmod(b, ldiv(l.K, l.B)) : b;
        else q = lmod(b, l.bpo);                  // each outer block is one inner call
        if (lmod(q, l.P) == 0) h += l.P;
    }
    if (l.kind == MS_FOR && lmod(b, l.bpo) == 0) {
        const uint64_t fb = ldiv(b, l.bpo);
        const uint64_t fq = l.K ? lmod(fb, ldiv(l.K, l.L)) : fb;
        if (lmod(fq, l.FP) == 0) h += 8 * l.FP;
    }
    return h;
}

// The values of the call holding value `first`: returns the array they are
// indexed in and sets `lo` to first's index in it (delta predecessors exist
// from index dstep of that array on).
__device__ __forceinline__ const uint64_t *chunk_src(const emit_layout &l, const uint64_t *v, uint64_t first,
                                                     uint64_t &lo)
{
    if (!l.cmap) { lo = first; return v; }
    const uint64_t c = first / l.K;
    lo = first - c * l.K;
    return v + (uint64_t)l.cmap[c] * l.K;
}

// Image offset of innermost block b's headers.
__device__ __forceinline__ uint64_t out_pos(const emit_layout &l, const uint64_t *start, uint64_t b)
{
    if (!l.cout) return start[b];
    const uint64_t bpc = ldiv(l.K, l.B), c = ldiv(b, bpc);
    return l.cout[c] + start[b] - start[c * bpc];
}

// xform over a value getter: get(gi) = value gi of the call's array.
template <class G>
__device__ __forceinline__ uint64_t xform_g(const emit_layout &l, const G &get, uint64_t gi, uint64_t ref)
{
    const uint64_t x = get(gi);
    if (l.kind == MS_DELTA) return x - (gi >= l.dstep ? get(gi - l.dstep) : 0ull);
    if (l.kind == MS_FOR) return x - ref;
    return x;
}

// ------------------------------------------------------------------
// Values straight from a select's flags (no materialised positions): the
// value of rank r is the row of the r-th selected row.
// ------------------------------------------------------------------
struct flag_src {
    sel_flags sf;
    const uint32_t *ustart;   // per innermost block b: the unit holding rank r0(b)
    uint32_t pred;            // delta predecessors a block reads (0 unless delta_f)
};

// First rank a block's kernels read: its predecessors included.
__device__ __forceinline__ uint64_t flag_r0(const emit_layout &l, const flag_src &fs, uint64_t b)
{
    const uint64_t first = b * l.B;
    return first >= fs.pred ? first - fs.pred : 0ull;
}

// One wave writes the rows of ranks [r0, r0 + cnt) to st(rank - r0, row),
// walking the units from u (a unit at or before the one holding rank r0)
// eight at a time.  Lane t takes rows 64 gl .. 64 gl + 63 of unit
// u + t / 8 (gl = t & 7): row 64 gl + 8 m + j of a unit is bit 8 gl + j of
// its ballot m (the count pass's thread mapping), so the lane's row word
// is byte gl of each of the unit's 8 ballots.  Its first rank is the
// unit's offset plus the set bits of the unit's lower row words (a scan
// over the lane's group of 8); then it writes its set bits in order.
template <class St>
__device__ __forceinline__ void gather_ranks(const sel_flags &sf, uint64_t nunits, uint64_t u, uint64_t r0,
                                             uint64_t cnt, const St &st)
{
    const uint32_t t = threadIdx.x & 63, gl = t & 7;
    const uint64_t r1 = r0 + cnt;
    if (cnt == 0 || cnt >= (1ull << 31)) return;   // callers ask for at most a few thousand ranks
    for (; u < nunits; u += 8) {
        const uint64_t uu = u + (t >> 3);
        const uint64_t ue = u + 8 < nunits ? u + 8 : nunits;
        const uint64_t end_rank = gload(sf.uoff + ue);   // requested with the group's loads
        uint64_t rw = 0, rank = 0;
        if (uu < nunits) {
            uint64_t b[8];
#pragma unroll
            for (int m = 0; m < 8; m++) b[m] = gload(sf.flags + 8 * uu + m);
            rank = gload(sf.uoff + uu);
            const uint32_t sh = 8 * gl;
#pragma unroll
            for (int m = 0; m < 8; m++) rw |= ((b[m] >> sh) & 0xffull) << (8 * m);
        }
        // exclusive scan of the row words' counts within the group of 8
        const uint32_t c = (uint32_t)__popcll(rw);
        uint32_t incl = c;
#pragma unroll
        for (uint32_t o = 1; o < 8; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o, 8);
            if (gl >= o) incl += y;
        }
        rank += incl - c;
        if (rank >= r1 || rank + c <= r0) rw = 0;
        // a lane with bits left has rank in (r0 - 64, r1): 32-bit from here
        int32_t rel = (int32_t)(int64_t)(rank - r0);
        const uint32_t n32 = (uint32_t)cnt;
        const uint64_t row0 = uu * 512 + 64 * gl;
        uint32_t x = (uint32_t)rw;
        while (x) {
            const uint32_t p = (uint32_t)__builtin_ctz(x);
            if ((uint32_t)rel < n32) st((uint64_t)rel, row0 + p);
            rel++;
            x &= x - 1;
        }
        x = (uint32_t)(rw >> 32);
        while (x) {
            const uint32_t p = (uint32_t)__builtin_ctz(x) + 32;
            if ((uint32_t)rel < n32) st((uint64_t)rel, row0 + p);
            rel++;
            x &= x - 1;
        }
        if (end_rank >= r1) break;
    }
}

// The select's split (thread 0): the selected rows below the scalar
// remainder's first row thr are counted from the unit offsets and thr's
// unit's ballots.
__device__ __forceinline__ void select_split(const sel_flags &sf, const sel_split &sp, uint64_t nunits)
{
    const ms_dev_meta *in = sp.in_meta;
    const ms_geom &g = sp.in_g;
    const ms_parts pt = ms_parts_of(g, in->count, in->size_used, in->size_compr);
    const bool has_rest = g.kind == MS_UNCOMPR ? in->count != 0 : in->size_compr != in->size_used;
    const uint64_t r = has_rest ? pt.n_rest % sp.lanes : 0;
    const uint64_t thr = pt.n - r;
    const uint64_t total = sf.uoff[nunits];
    const uint64_t u = thr / 512, w = thr % 512;
    uint64_t na = u < nunits ? sf.uoff[u] : total;
    if (u < nunits && w) {
        // row 64 gl + 8 m + j of the unit is bit 8 gl + j of ballot m
        for (uint32_t m = 0; m < 8; m++) {
            const uint64_t B = sf.flags[8 * u + m];
            for (uint32_t gl = 0; gl < 8; gl++) {
                const uint64_t row = 64 * gl + 8 * m;
                if (row >= w) continue;
                const uint64_t k = w - row < 8 ? w - row : 8;
                na += __popcll((B >> (8 * gl)) & ((1ull << k) - 1));
            }
        }
    }
    *sp.d_cnt = total;
    *sp.d_na = na;
    *sp.d_flag = r != 0;
}

// ustart[b] for every block whose first read rank r0(b) lies in unit u's
// ranks [uoff[u], uoff[u+1]); one thread per unit.  Thread 0 also makes
// the select's split.
__global__ __launch_bounds__(WG) void k_flag_starts(emit_layout l, flag_src fs, sel_split sp, uint64_t cap,
                                                    uint32_t *ustart, ms_dev_meta *om)
{
    const uint64_t nunits = *fs.sf.d_units;
    if (blockIdx.x == 0 && threadIdx.x == 0) select_split(fs.sf, sp, nunits);
    // the output's meta starts cleared: the first kernel of the emission
    if (blockIdx.x == 0 && threadIdx.x < sizeof(ms_dev_meta) / 8) ((uint64_t *)om)[threadIdx.x] = 0;
    for (uint64_t u = (uint64_t)blockIdx.x * WG + threadIdx.x; u < nunits; u += (uint64_t)gridDim.x * WG) {
        const uint64_t lo = fs.sf.uoff[u], hi = fs.sf.uoff[u + 1];
        if (lo == hi) continue;
        // r0(b) = max(b B - pred, 0): b = 0 starts at rank 0
        uint64_t b = (lo + fs.pred + l.B - 1) / l.B;
        if (lo == 0) {
            if (cap > 0) ustart[0] = (uint32_t)u;
            if (b == 0) b = 1;
        }
        for (; b < cap && b * l.B - fs.pred < hi; b++) ustart[b] = (uint32_t)u;
    }
}

// for.h:121-161 reference value of one outer block, one wave per block.
// FL: the values come from a select's flags (dynamic LDS: L words).
template <bool FL>
__global__ __launch_bounds__(64) void k_for_refs(emit_layout l, const uint64_t *v, flag_src fs, const uint64_t *d_m,
                                                 uint64_t *refs)
{
    extern __shared__ unsigned long long fvals[];
    const emit_plan pl = plan_of(l, *d_m);
    const uint32_t t = threadIdx.x;
    const uint64_t nunits = FL ? *fs.sf.d_units : 0;
    for (uint64_t fb = blockIdx.x; fb < pl.nfb; fb += gridDim.x) {
        uint64_t lo;
        const uint64_t *xg = nullptr;
        if (FL) {
            gather_ranks(fs.sf, nunits, fs.ustart[fb * l.L / l.B], fb * l.L, l.L,
                         [&](uint64_t i, uint64_t row) { fvals[i] = row; });
            __syncthreads();
        } else {
            xg = chunk_src(l, v, fb * l.L, lo) + lo;
        }
        const auto x = [&](uint64_t i) -> uint64_t { return FL ? (uint64_t)fvals[i] : xg[i]; };
        uint64_t r;
        if (l.E == 1) {
            r = ~0ull;
            for (uint64_t i = t; i < l.L; i += 64) r = x(i) < r ? x(i) : r;
            for (int o = 32; o >= 1; o >>= 1) {
                const uint64_t y = __shfl_xor(r, o, 64);
                r = y < r ? y : r;
            }
        } else {
            // lane j = i % E: signed minimum per lane, then unsigned over lanes
            int64_t sm = INT64_MAX;
            for (uint64_t i = t; i < l.L; i += 64) sm = (int64_t)x(i) < sm ? (int64_t)x(i) : sm;
            for (uint32_t o = 32; o >= l.E; o >>= 1) {
                const int64_t y = (int64_t)__shfl_xor((uint64_t)sm, o, 64);
                sm = y < sm ? y : sm;
            }
            r = (uint64_t)sm;
            for (uint32_t o = l.E >> 1; o >= 1; o >>= 1) {
                const uint64_t y = __shfl_xor(r, o, 64);
                r = y < r ? y : r;
            }
        }
        if (t == 0) refs[fb] = r == ~0ull ? r - 1 : r;
        if (FL) __syncthreads();
    }
}

// Bit width of innermost block b over a value getter (one wave; lane t).
template <class G>
__device__ __forceinline__ uint32_t block_bw(const emit_layout &l, const G &get, uint64_t lo, uint64_t ref,
                                             uint32_t t)
{
    uint64_t o = 1;
    for (uint64_t i = t; i < l.B; i += 64) o |= xform_g(l, get, lo + i, ref);
    o = wave_or_u64(o);
    return 64 - __clzll(o);
}

// Assembles the interleaved words of innermost block b (one wave, lane t)
// in lds (B words; the first B bw / 64 hold the block).  sync() orders the
// wave's LDS accesses (the workgroup barrier when the workgroup is this
// wave); the words are complete when it returns.
template <class G, class Sync>
__device__ __forceinline__ void assemble_block(const emit_layout &l, uint32_t bw, const G &get, uint64_t lo,
                                               uint64_t ref, unsigned long long *lds, uint32_t t, const Sync &sync)
{
    // Lane counts that divide the wave and blocks of whole waves of codes
    // let each thread own a run of consecutive codes of one lane.
    const uint32_t S = (uint32_t)l.S;
    const bool runs = S != 0 && S <= 64 && 64 % S == 0 && l.B % 64 == 0;
    const uint32_t j = runs ? t % S : 0, g = runs ? t / S : 0, C = (uint32_t)(l.B >> 6);
    const uint64_t words = l.B * bw / 64;
    for (uint64_t w = t; w < words; w += 64) lds[w] = 0;
    sync();
    // OR every code in, unmasked (vbp_routine_gen.py:101, 111)
    if (runs) {
        // Thread (j, g) owns codes g*C .. g*C+C-1 of lane j: consecutive
        // codes share words, so it ORs them in registers and issues one
        // LDS atomic per word it touches instead of one per code.
        const uint64_t base = lo + j;
        uint32_t k = g * C, bit = k * bw, cw = bit >> 6;
        uint64_t acc = 0;
        // A delta step of one lane width makes each code's predecessor
        // the thread's previous code: one extra load per run.
        const bool own_pred = l.kind == MS_DELTA && l.dstep == S;
        uint64_t prev = 0;
        if (own_pred) {
            const uint64_t g0 = base + (uint64_t)k * S;
            prev = g0 >= S ? get(g0 - S) : 0ull;
        }
        for (uint32_t c0 = 0; c0 < C; c0 += 8) {
            uint64_t x[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                const uint64_t gi = base + (uint64_t)(k + q) * S;
                x[q] = c0 + q < C ? (own_pred ? get(gi) : xform_g(l, get, gi, ref)) : 0ull;
            }
            if (own_pred) {
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q) {
                    const uint64_t r = x[q];
                    x[q] = r - prev;
                    prev = r;
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                if (c0 + q >= C) break;
                const uint32_t w = bit >> 6, sh = bit & 63;
                if (w != cw) {
                    atomicOr(&lds[cw * S + j], (unsigned long long)acc);
                    acc = 0;
                    cw = w;
                }
                acc |= x[q] << sh;
                if (sh + bw > 64) {
                    atomicOr(&lds[cw * S + j], (unsigned long long)acc);
                    cw = w + 1;
                    acc = x[q] >> (64 - sh);
                }
                bit += bw;
            }
            k += 8;
        }
        atomicOr(&lds[cw * S + j], (unsigned long long)acc);
    } else {
        for (uint64_t i = t; i < l.B; i += 64) {
            const uint64_t c = xform_g(l, get, lo + i, ref);
            const uint64_t j = i % l.S, k = i / l.S;
            const uint64_t bit = k * bw;
            const uint64_t w = bit >> 6;
            const uint32_t sh = (uint32_t)(bit & 63);
            atomicOr(&lds[w * l.S + j], (unsigned long long)(c << sh));
            if (sh + bw > 64) atomicOr(&lds[(w + 1) * l.S + j], (unsigned long long)(c >> (64 - sh)));
        }
    }
    sync();
}

// What follows block b's words at image offset off: the block index entry,
// the dvbp page meta and the for_f reference page.  Lanes t < NL of a group
// of NL lanes write them.
template <uint32_t NL = 64>
__device__ __forceinline__ void finish_block(const emit_layout &l, const emit_plan &pl, uint64_t b, uint32_t bw,
                                             uint64_t off, uint64_t ref, const uint64_t *refs, const uint64_t *start,
                                             uint8_t *out, uint64_t *index, uint64_t *refs_out, uint32_t t)
{
    const uint64_t fb = ldiv(b, l.bpo);
    if (t == 0) index[b] = (off << 8) | bw;
    // dynamic_vbp page meta (dynamic_vbp.h:216-255)
    if (l.leaf == MS_DYNAMIC_VBP) {
        uint64_t q, call_end;
        if (l.kind == MS_DYNAMIC_VBP) {
            if (l.K) {
                const uint64_t bpc = ldiv(l.K, l.B);
                q = lmod(b, bpc);
                call_end = b - q + bpc;
                if (call_end > pl.nb) call_end = pl.nb;
            } else {
                q = b;
                call_end = pl.nb;
            }
        } else {
            q = lmod(b, l.bpo);
            call_end = b - q + l.bpo;
        }
        const uint64_t qp = lmod(q, l.P);
        const uint64_t bpf = b - qp;                                      // first block of the page
        const uint64_t meta = out_pos(l, start, bpf) + hdr_of(l, bpf) - l.P; // meta bytes end the header
        if (t == 0) out[meta + qp] = (uint8_t)bw;
        if (b == bpf) {
            const uint64_t npage = call_end - bpf < l.P ? call_end - bpf : l.P;
            for (uint64_t x = npage + t; x < l.P; x += NL) out[meta + x] = (uint8_t)MS_VBP_BW_NOBLOCK;
        }
    }
    // for_f reference page (for.h:207-233)
    if (l.kind == MS_FOR && lmod(b, l.bpo) == 0) {
        if (t == 0) refs_out[fb] = ref;
        const uint64_t fpc = l.K ? ldiv(l.K, l.L) : 0;
        const uint64_t fq = l.K ? lmod(fb, fpc) : fb;
        if (lmod(fq, l.FP) == 0) {
            uint64_t fcall_end = l.K ? fb - lmod(fb, fpc) + fpc : pl.nfb;
            if (fcall_end > pl.nfb) fcall_end = pl.nfb;
            const uint64_t nfp = fcall_end - fb < l.FP ? fcall_end - fb : l.FP;
            uint64_t *rp = (uint64_t *)(out + out_pos(l, start, b));
            for (uint64_t x = t; x < l.FP; x += NL) rp[x] = x < nfp ? refs[fb + x] : (x == nfp ? ~0ull : 0ull);
        }
    }
}

// Packs innermost block b (one wave, lane t) at image offset off: its words
// are assembled in lds (B words) and written out, then finish_block.
template <class G, class Sync>
__device__ __forceinline__ void pack_block(const emit_layout &l, const emit_plan &pl, uint64_t b, uint32_t bw,
                                           uint64_t off, const G &get, uint64_t lo, uint64_t ref, const uint64_t *refs,
                                           const uint64_t *start, unsigned long long *lds, uint8_t *out,
                                           uint64_t *index, uint64_t *refs_out, uint32_t t, const Sync &sync)
{
    assemble_block(l, bw, get, lo, ref, lds, t, sync);
    const uint64_t words = l.B * bw / 64;
    uint64_t *dst = (uint64_t *)(out + off);
    for (uint64_t w = t; w < words; w += 64) dst[w] = lds[w];
    sync();
    finish_block(l, pl, b, bw, off, ref, refs, start, out, index, refs_out, t);
}

// Bit width and header+payload bytes of every innermost block.
// With d_nb the blocks past the column's own (*d_nb, written here) are left
// alone, for a scan sized by that count; without it they get size 0.
__global__ __launch_bounds__(64) void k_block_sizes(emit_layout l, const uint64_t *v, const uint64_t *d_m,
                                                    const uint64_t *refs, uint64_t nb_bound, uint64_t *sz,
                                                    uint32_t *bws, uint64_t *d_nb)
{
    const emit_plan pl = plan_of(l, *d_m);
    const uint32_t t = threadIdx.x;
    if (d_nb) {
        if (blockIdx.x == 0 && t == 0) *d_nb = pl.nb;
        if (nb_bound > pl.nb) nb_bound = pl.nb;
    }
    for (uint64_t b = blockIdx.x; b < nb_bound; b += gridDim.x) {
        if (b >= pl.nb) {
            if (t == 0) { sz[b] = 0; bws[b] = 0; }
            continue;
        }
        uint32_t bw = l.bw;
        if (l.leaf == MS_DYNAMIC_VBP) {
            const uint64_t ref = l.kind == MS_FOR ? refs[b / l.bpo] : 0;
            uint64_t lo;
            const uint64_t *src = chunk_src(l, v, b * l.B, lo);
            bw = block_bw(l, [&](uint64_t gi) { return src[gi]; }, lo, ref, t);
        }
        if (t == 0) {
            sz[b] = hdr_of(l, b) + l.B * bw / 8;
            bws[b] = bw;
        }
    }
}

// Packs one innermost block per wave.  Dynamic LDS: B 64-bit words.
__global__ __launch_bounds__(64) void k_pack_blocks(emit_layout l, const uint64_t *v, const uint64_t *d_m,
                                                    const uint64_t *refs, const uint64_t *start, const uint32_t *bws,
                                                    uint8_t *out, uint64_t *index, uint64_t *refs_out)
{
    extern __shared__ unsigned long long lds[];
    const emit_plan pl = plan_of(l, *d_m);
    const uint32_t t = threadIdx.x;
    for (uint64_t b = blockIdx.x; b < pl.nb; b += gridDim.x) {
        const uint32_t bw = bws[b];
        const uint64_t off = out_pos(l, start, b) + hdr_of(l, b);
        if (off + 8 * (l.B * bw / 64) > l.cap) {   // uniform per wave
            if (t == 0) emit_overflow(l);
            continue;
        }
        uint64_t lo;
        const uint64_t *src = chunk_src(l, v, b * l.B, lo);
        const uint64_t ref = l.kind == MS_FOR ? refs[b / l.bpo] : 0;
        pack_block(l, pl, b, bw, off, [&](uint64_t gi) { return src[gi]; }, lo, ref, refs, start, lds, out, index,
                   refs_out, t, [] { __syncthreads(); });
    }
}

// Sizes and packed words from a select's flags, one innermost block per
// wave (dynamic LDS: B packed words, then the B + pred gathered values).
// Block b's words go to stage + b B, its size to sz[b], its width to
// bws[b]; *d_nb receives the number of blocks (the scan's length).  The
// positions read by rank later go to pos: a delta tile's predecessors
// (gathered with the block the tile starts at) and, by workgroup 0 after
// its blocks, the tail behind the compressed part.
__global__ __launch_bounds__(64) void k_flag_stage(emit_layout l, flag_src fs, const uint64_t *d_m,
                                                   const uint64_t *refs, const uint64_t *d_total, uint64_t *d_nb,
                                                   uint64_t *sz, uint32_t *bws, uint64_t *stage, uint64_t *pos)
{
    extern __shared__ unsigned long long lds[];
    unsigned long long *fvals = lds + l.B;
    const emit_plan pl = plan_of(l, *d_m);
    const uint32_t t = threadIdx.x;
    const uint64_t nunits = *fs.sf.d_units;
    if (blockIdx.x == 0 && t == 0) *d_nb = pl.nb;
    for (uint64_t b = blockIdx.x; b < pl.nb; b += gridDim.x) {
        const uint64_t r0 = flag_r0(l, fs, b), lo = b * l.B;
        gather_ranks(fs.sf, nunits, fs.ustart[b], r0, lo + l.B - r0,
                     [&](uint64_t i, uint64_t row) { fvals[i] = row; });
        __syncthreads();
        if (fs.pred && lo > 0 && lo % MS_DELTA_TILE == 0)
            for (uint32_t i = t; i < fs.pred; i += 64) pos[r0 + i] = fvals[i];
        const uint64_t ref = l.kind == MS_FOR ? refs[b / l.bpo] : 0;
        const auto get = [&](uint64_t gi) -> uint64_t { return fvals[gi - r0]; };
        const uint32_t bw = block_bw(l, get, lo, ref, t);
        if (t == 0) {
            sz[b] = hdr_of(l, b) + l.B * bw / 8;
            bws[b] = bw;
        }
        assemble_block(l, bw, get, lo, ref, lds, t, [] { __syncthreads(); });
        const uint64_t words = l.B * bw / 64;
        uint64_t *dst = stage + b * l.B;
        for (uint64_t w = t; w < words; w += 64) dst[w] = lds[w];
        __syncthreads();
    }
    if (blockIdx.x == 0) {
        const uint64_t total = *d_total, r0 = pl.m_c;
        const uint64_t cnt = total > r0 ? total - r0 : 0;
        if (cnt) gather_ranks(fs.sf, nunits, fs.ustart[pl.nb], r0, cnt,
                              [&](uint64_t i, uint64_t row) { pos[r0 + i] = row; });
    }
}

// Moves each staged block to its offset in the image and writes what
// follows it (index entry, page meta, reference page).  A group of 16
// lanes takes a block, so a wave has four blocks' loads in flight; four
// waves per workgroup.
__global__ __launch_bounds__(256) void k_flag_place(emit_layout l, const uint64_t *d_m, const uint64_t *refs,
                                                    const uint64_t *start, const uint32_t *bws,
                                                    const uint64_t *stage, uint8_t *out, uint64_t *index,
                                                    uint64_t *refs_out)
{
    constexpr uint32_t NL = 16;
    const emit_plan pl = plan_of(l, *d_m);
    const uint32_t t = threadIdx.x % NL;
    const uint64_t step = (uint64_t)gridDim.x * (256 / NL);
    for (uint64_t b = (uint64_t)blockIdx.x * (256 / NL) + threadIdx.x / NL; b < pl.nb; b += step) {
        const uint64_t *src = stage + b * l.B;
        const uint32_t bw = bws[b];
        const uint64_t off = out_pos(l, start, b) + hdr_of(l, b);
        const uint64_t words = l.B * bw / 64;
        if (off + 8 * words > l.cap) {
            if (t == 0) emit_overflow(l);
            continue;
        }
        uint64_t *dst = (uint64_t *)(out + off);
        for (uint64_t w = t; w < words; w += NL) dst[w] = src[w];
        const uint64_t ref = l.kind == MS_FOR ? refs[ldiv(b, l.bpo)] : 0;
        finish_block<NL>(l, pl, b, bw, off, ref, refs, start, out, index, refs_out, t);
    }
}

// Static bit width: every output word straight from its codes.
__global__ __launch_bounds__(WG) void k_pack_static(emit_layout l, const uint64_t *v, const uint64_t *d_m,
                                                    uint8_t *out)
{
    const emit_plan pl = plan_of(l, *d_m);
    const uint64_t bw = l.bw, S = l.S;
    const uint64_t nwords = pl.m_c * bw / 64;
    if (nwords * 8 > l.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) emit_overflow(l);
        return;
    }
    uint64_t *o = (uint64_t *)out;
    for (uint64_t W = (uint64_t)blockIdx.x * WG + threadIdx.x; W < nwords; W += (uint64_t)gridDim.x * WG) {
        const uint64_t j = W % S, w = W / S;
        const uint64_t b0 = 64 * w, b1 = 64 * w + 64;
        uint64_t k = (b0 + bw - 1) / bw;          // first code starting in this word
        uint64_t word = 0;
        if (k > 0) {
            const uint64_t kc = k - 1;            // code crossing in from the previous word
            if (kc * bw + bw > b0) word |= v[kc * S + j] >> (b0 - kc * bw);
        }
        for (; k * bw < b1; k++) word |= v[k * S + j] << (k * bw - b0);
        o[W] = word;
    }
}

// The end of an emission in one launch: the output's meta (block 0), the
// delta_f tile carries (carries non-null) and the raw rest behind the gap;
// every workgroup derives the gap's end itself.  start / bws are null for
// static_vbp_f.
// With d_flag (a select's split): values [*d_m, *d_total) are appended raw
// behind the output when *d_flag (select_compr.h:209-250).
__global__ __launch_bounds__(WG) void k_emit_tail(emit_layout l, const uint64_t *v, const uint64_t *d_m,
                                                  const uint64_t *start, const uint32_t *bws, ms_dev_meta *meta,
                                                  uint64_t *d_append, uint64_t *carries, uint8_t *out,
                                                  const uint64_t *d_total, const uint64_t *d_flag)
{
    const emit_plan pl = plan_of(l, *d_m);
    uint64_t C;
    if (l.kind == MS_STATIC_VBP) C = pl.m_c * l.bw / 8;
    else if (pl.nb == 0) C = 0;
    else C = start[pl.nb - 1] + hdr_of(l, pl.nb - 1) + l.B * bws[pl.nb - 1] / 8;
    const uint64_t R = ms_align64(C + l.gap);
    const uint64_t rest = pl.m - pl.m_c;
    const bool fits = R + 8 * rest <= l.cap;
    const uint64_t tid = (uint64_t)blockIdx.x * WG + threadIdx.x, nth = (uint64_t)gridDim.x * WG;
    const uint64_t app = rest ? R + 8 * rest : R;   // the write-iterator's append offset
    const bool appending = d_flag && *d_flag;
    const uint64_t total = appending ? *d_total : pl.m;
    const bool afits = !appending || app + 8 * (total - pl.m) <= l.cap;
    if (tid == 0) {
        if (!fits || !afits) emit_overflow(l);
        const bool app_done = appending && afits;
        meta->count = app_done ? total : pl.m;
        meta->size_compr = C;
        meta->size_used = app_done ? app + 8 * (total - pl.m) : rest ? R + 8 * rest : C;
        meta->aux[0] = app;
        meta->aux[1] = R;
        if (d_append) *d_append = app;
    }
    if (appending && afits) {
        uint64_t *dst = (uint64_t *)(out + app);
        for (uint64_t i = pl.m + tid; i < total; i += nth) dst[i - pl.m] = v[i];
    }
    if (carries) {
        const uint64_t ntiles = (pl.m_c + MS_DELTA_TILE - 1) / MS_DELTA_TILE;
        for (uint64_t i = tid; i < ntiles * 8; i += nth) {
            const uint64_t tile = i / 8, j = i % 8;
            carries[i] = (tile == 0 || j >= l.dstep) ? 0ull : v[tile * MS_DELTA_TILE - l.dstep + j];
        }
    }
    if (fits) {
        uint64_t *dst = (uint64_t *)(out + R);
        for (uint64_t i = tid; i < rest; i += nth) dst[i] = v[pl.m_c + i];
    }
}

// The emit layout of an output column written as `chunk`-value calls.
static emit_layout out_layout(ms_col *out, uint64_t chunk)
{
    const ms_format_desc &f = out->f;
    const ms_geom &g = out->g;
    emit_layout l = {};
    l.kind = f.kind;
    l.leaf = g.leaf;
    l.bw = g.bw;
    l.S = g.lanes;
    l.dstep = g.dstep;
    l.B = g.inner_bs;
    l.P = g.leaf_page;
    l.bs = g.bs;
    l.gap = g.gap;
    l.K = chunk;
    l.cap = out->size_alloc;
    l.status = &out->meta->status;
    if (f.kind == MS_DELTA || f.kind == MS_FOR) {
        l.L = g.outer_block;
        l.FP = g.outer_page;
        l.bpo = g.outer_block / g.inner_bs;
        l.E = f.kind == MS_FOR ? (f.step ? f.step : g.lanes) : 1;
    } else {
        l.L = l.B;
        l.bpo = 1;
    }
    return l;
}

int launch_emit(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, uint64_t chunk,
                uint64_t *d_append, temps &sc)
{
    hipStream_t s = sc.s;
    const ms_format_desc &f = out->f;
    const emit_layout l = out_layout(out, chunk);
    if (f.kind != MS_STATIC_VBP && l.B * 8 > 65536)
        return ms_fail(MS_ERR_INVALID, "emit: innermost blocks above 8192 values are not supported");
    const uint64_t nb_bound = m_bound / l.B + 1;
    if (f.kind == MS_STATIC_VBP) {
        hipLaunchKernelGGL(k_pack_static, dim3(grid_for(m_bound * l.bw / 64 + 1, WG * 4ull)), dim3(WG), 0, s, l, d_v,
                           d_m, out->data);
        MS_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_emit_tail, dim3(grid_for(l.bs + l.K + 1, WG, 64)), dim3(WG), 0, s, l, d_v, d_m,
                           (const uint64_t *)nullptr, (const uint32_t *)nullptr, out->meta, d_append,
                           (uint64_t *)nullptr, out->data, (const uint64_t *)nullptr, (const uint64_t *)nullptr);
    } else {
        uint64_t *sz = (uint64_t *)sc.get(nb_bound * 8 + 64);
        uint64_t *start = (uint64_t *)sc.get(nb_bound * 8 + 64);
        uint32_t *bws = (uint32_t *)sc.get(nb_bound * 4 + 64);
        uint64_t *refs = nullptr;
        if (!sz || !start || !bws) return ms_fail(MS_ERR_NOMEM, "emit scratch");
        if (f.kind == MS_FOR) {
            refs = (uint64_t *)sc.get((m_bound / l.L + 1) * 8 + 64);
            if (!refs) return ms_fail(MS_ERR_NOMEM, "emit scratch");
            hipLaunchKernelGGL(k_for_refs<false>, dim3(grid_for(m_bound / l.L + 1, 1, 65536)), dim3(64), 0, s, l, d_v,
                               flag_src{}, d_m, refs);
            MS_HIP(hipGetLastError());
        }
        uint64_t *d_nb = (uint64_t *)sc.get(64);
        if (!d_nb) return ms_fail(MS_ERR_NOMEM, "emit scratch");
        hipLaunchKernelGGL(k_block_sizes, dim3(grid_for(nb_bound, 1, 65536)), dim3(64), 0, s, l, d_v, d_m,
                           (const uint64_t *)refs, nb_bound, sz, bws, d_nb);
        MS_HIP(hipGetLastError());
        MS_TRY(launch_exclusive_scan_n(sz, start, d_nb, 0, nb_bound, sc));
        if (out->index_cap < nb_bound - 1 || (f.kind == MS_FOR && out->refs_cap < m_bound / l.L))
            return ms_fail(MS_ERR_CAPACITY, "emit: index capacity");
        hipLaunchKernelGGL(k_pack_blocks, dim3(grid_for(nb_bound, 1, 65536)), dim3(64), l.B * 8, s, l, d_v, d_m,
                           (const uint64_t *)refs, (const uint64_t *)start, (const uint32_t *)bws, out->data,
                           out->index, out->refs);
        MS_HIP(hipGetLastError());
        if (f.kind == MS_DELTA && out->refs_cap < (m_bound / MS_DELTA_TILE + 1) * 8)
            return ms_fail(MS_ERR_CAPACITY, "emit: delta carries capacity");
        const uint64_t tail = std::max<uint64_t>(l.bs + l.K + 1, m_bound / MS_DELTA_TILE * 8 + 8);
        hipLaunchKernelGGL(k_emit_tail, dim3(grid_for(tail, WG, 1024)), dim3(WG), 0, s, l, d_v, d_m,
                           (const uint64_t *)start, (const uint32_t *)bws, out->meta, d_append,
                           f.kind == MS_DELTA ? out->refs : (uint64_t *)nullptr, out->data, (const uint64_t *)nullptr,
                           (const uint64_t *)nullptr);
    }
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------
// Emission straight from a select's flags.
// ------------------------------------------------------------------
bool emit_flags_ok(const ms_format_desc &f)
{
    if (f.kind != MS_DYNAMIC_VBP && f.kind != MS_DELTA && f.kind != MS_FOR) return false;
    const ms_geom g = ms_make_geom(f);
    if (g.leaf != MS_DYNAMIC_VBP || g.inner_bs == 0 || g.inner_bs > 2048 || 2048 % g.inner_bs) return false;
    if (f.kind == MS_DELTA && (g.dstep == 0 || g.dstep > 64 || g.dstep >= g.inner_bs)) return false;
    if (f.kind == MS_FOR && g.outer_block > 4096) return false;
    return true;
}

int launch_emit_flags(ms_col *out, const sel_flags &sf, const sel_split &sp, uint64_t *pos, uint64_t m_bound,
                      uint64_t chunk, temps &sc)
{
    const uint64_t *d_m = sp.d_na, *d_total = sp.d_cnt;
    hipStream_t s = sc.s;
    const ms_format_desc &f = out->f;
    if (!emit_flags_ok(f)) return ms_fail(MS_ERR_INVALID, "emit from flags: format kind %u", f.kind);
    const emit_layout l = out_layout(out, chunk);
    const uint64_t nb_bound = m_bound / l.B + 1;
    flag_src fs = {};
    fs.sf = sf;
    fs.pred = f.kind == MS_DELTA ? l.dstep : 0;
    // blocks whose first read rank exists: b B - pred < m_bound
    const uint64_t starts = (m_bound + fs.pred) / l.B + 2;
    uint32_t *ustart = (uint32_t *)sc.get(starts * 4 + 64);
    uint64_t *sz = (uint64_t *)sc.get(nb_bound * 8 + 64);
    uint64_t *start = (uint64_t *)sc.get(nb_bound * 8 + 64);
    uint32_t *bws = (uint32_t *)sc.get(nb_bound * 4 + 64);
    if (!ustart || !sz || !start || !bws) return ms_fail(MS_ERR_NOMEM, "emit scratch");
    if (sf.units_bound >= (1ull << 32)) return ms_fail(MS_ERR_CAPACITY, "emit from flags: %llu units",
                                                      (unsigned long long)sf.units_bound);
    fs.ustart = ustart;
    hipLaunchKernelGGL(k_flag_starts, dim3(grid_for(sf.units_bound, WG)), dim3(WG), 0, s, l, fs, sp, starts, ustart,
                       out->meta);
    MS_HIP(hipGetLastError());
    uint64_t *refs = nullptr;
    if (f.kind == MS_FOR) {
        refs = (uint64_t *)sc.get((m_bound / l.L + 1) * 8 + 64);
        if (!refs) return ms_fail(MS_ERR_NOMEM, "emit scratch");
        hipLaunchKernelGGL(k_for_refs<true>, dim3(grid_for(m_bound / l.L + 1, 1, 65536)), dim3(64), l.L * 8, s, l,
                           (const uint64_t *)nullptr, fs, d_m, refs);
        MS_HIP(hipGetLastError());
    }
    const uint32_t vals_lds = (uint32_t)((l.B + fs.pred) * 8);
    const uint32_t fl_grid = grid_for(nb_bound, 1, 65536);
    uint64_t *d_nb = (uint64_t *)sc.get(64);
    if (!d_nb) return ms_fail(MS_ERR_NOMEM, "emit scratch");
    // every block's words staged at a fixed slot of B words, so the packing
    // happens in the same pass as the sizes and placing them is a copy
    uint64_t *stage = (uint64_t *)sc.get(nb_bound * l.B * 8 + 64);
    if (!stage) return ms_fail(MS_ERR_NOMEM, "emit scratch");
    hipLaunchKernelGGL(k_flag_stage, dim3(fl_grid), dim3(64), l.B * 8 + vals_lds, s, l, fs, d_m,
                       (const uint64_t *)refs, d_total, d_nb, sz, bws, stage, pos);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan_n(sz, start, d_nb, 0, nb_bound, sc));
    if (out->index_cap < nb_bound - 1 || (f.kind == MS_FOR && out->refs_cap < m_bound / l.L))
        return ms_fail(MS_ERR_CAPACITY, "emit: index capacity");
    hipLaunchKernelGGL(k_flag_place, dim3(grid_for(nb_bound, 16, 16384)), dim3(256), 0, s, l, d_m,
                       (const uint64_t *)refs, (const uint64_t *)start, (const uint32_t *)bws,
                       (const uint64_t *)stage, out->data, out->index, out->refs);
    MS_HIP(hipGetLastError());
    if (f.kind == MS_DELTA && out->refs_cap < (m_bound / MS_DELTA_TILE + 1) * 8)
        return ms_fail(MS_ERR_CAPACITY, "emit: delta carries capacity");
    const uint64_t tail = std::max<uint64_t>(l.bs + l.K + 1, m_bound / MS_DELTA_TILE * 8 + 8);
    hipLaunchKernelGGL(k_emit_tail, dim3(grid_for(tail, WG, 1024)), dim3(WG), 0, s, l, (const uint64_t *)pos, d_m,
                       (const uint64_t *)start, (const uint32_t *)bws, out->meta, (uint64_t *)nullptr,
                       f.kind == MS_DELTA ? out->refs : (uint64_t *)nullptr, out->data, d_total,
                       (const uint64_t *)sp.d_flag);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------
// Chunked emission: one morph_batch call per chunk of K values.
// ------------------------------------------------------------------
static emit_layout chunk_layout(const chunk_emit &ce)
{
    const ms_format_desc &f = ce.f;
    const ms_geom g = ms_make_geom(f);
    emit_layout l = {};
    l.kind = f.kind;
    l.leaf = g.leaf;
    l.bw = g.bw;
    l.S = g.lanes;
    l.dstep = g.dstep;
    l.B = g.inner_bs;
    l.P = g.leaf_page;
    l.bs = g.bs;
    l.K = ce.K;
    l.cap = ~0ull;
    if (f.kind == MS_DELTA || f.kind == MS_FOR) {
        l.L = g.outer_block;
        l.FP = g.outer_page;
        l.bpo = g.outer_block / g.inner_bs;
        l.E = f.kind == MS_FOR ? (f.step ? f.step : g.lanes) : 1;
    } else {
        l.L = l.B;
        l.bpo = 1;
    }
    l.cmap = ce.cmap;
    return l;
}

__global__ __launch_bounds__(WG) void k_chunk_bytes(const uint64_t *start, uint64_t bpc, uint64_t n, uint64_t *csz)
{
    for (uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x; c < n; c += (uint64_t)gridDim.x * WG)
        csz[c] = start[(c + 1) * bpc] - start[c * bpc];
}

int launch_chunk_sizes(chunk_emit &ce, uint64_t *d_csz, temps &sc)
{
    hipStream_t s = sc.s;
    const emit_layout l = chunk_layout(ce);
    const ms_geom g = ms_make_geom(ce.f);
    if (g.leaf != MS_DYNAMIC_VBP || (ce.f.kind != MS_DYNAMIC_VBP && ce.f.kind != MS_DELTA && ce.f.kind != MS_FOR))
        return ms_fail(MS_ERR_INVALID, "chunked emission: dynamic_vbp_f or a cascade over it");
    if (ce.K == 0 || ce.K % l.L || ce.K % l.B || l.B * 8 > 65536)
        return ms_fail(MS_ERR_INVALID, "chunked emission: chunk of %llu values", (unsigned long long)ce.K);
    const uint64_t m = ce.nchunks * ce.K;
    ce.nb = m / l.B;
    ce.d_m = (uint64_t *)sc.get(64);
    ce.sz = (uint64_t *)sc.get((ce.nb + 1) * 8 + 64);
    ce.start = (uint64_t *)sc.get((ce.nb + 1) * 8 + 64);
    ce.bws = (uint32_t *)sc.get((ce.nb + 1) * 4 + 64);
    ce.refs = nullptr;
    if (!ce.d_m || !ce.sz || !ce.start || !ce.bws) return ms_fail(MS_ERR_NOMEM, "chunked emission scratch");
    MS_TRY(launch_fill_u64(ce.d_m, 1, m, s));
    if (ce.nchunks == 0) return MS_OK;
    if (ce.f.kind == MS_FOR) {
        ce.refs = (uint64_t *)sc.get((m / l.L + 1) * 8 + 64);
        if (!ce.refs) return ms_fail(MS_ERR_NOMEM, "chunked emission scratch");
        hipLaunchKernelGGL(k_for_refs<false>, dim3(grid_for(m / l.L, 1, 65536)), dim3(64), 0, s, l, ce.v, flag_src{},
                           ce.d_m, ce.refs);
        MS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_block_sizes, dim3(grid_for(ce.nb + 1, 1, 65536)), dim3(64), 0, s, l, ce.v, ce.d_m,
                       (const uint64_t *)ce.refs, ce.nb + 1, ce.sz, ce.bws, (uint64_t *)nullptr);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan(ce.sz, ce.start, ce.nb + 1, sc));
    if (d_csz) {
        hipLaunchKernelGGL(k_chunk_bytes, dim3(grid_for(ce.nchunks, WG)), dim3(WG), 0, s, (const uint64_t *)ce.start,
                           ce.K / l.B, ce.nchunks, d_csz);
        MS_HIP(hipGetLastError());
    }
    return MS_OK;
}

int launch_chunk_pack(const chunk_emit &ce, const uint64_t *d_cout, uint8_t *out, uint64_t cap, uint64_t *status,
                      uint64_t *index_out, uint64_t *refs_out, hipStream_t s)
{
    if (ce.nchunks == 0) return MS_OK;
    emit_layout l = chunk_layout(ce);
    l.cout = d_cout;
    l.cap = cap;
    l.status = status;
    hipLaunchKernelGGL(k_pack_blocks, dim3(grid_for(ce.nb, 1, 65536)), dim3(64), l.B * 8, s, l, ce.v,
                       (const uint64_t *)ce.d_m, (const uint64_t *)ce.refs, (const uint64_t *)ce.start,
                       (const uint32_t *)ce.bws, out, index_out, refs_out);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------
// rle_f (rle.h:95-147): maximal runs.  The column is cut into chunks of
// RLE_CCH values, a wave per chunk.  A run belongs to the chunk holding
// its first value; pass 1 counts each chunk's run starts (ballots) and
// records its first start, a scan of the counts places every chunk's
// runs, pass 2 writes each run's value and, when the next start lies in
// the same chunk, its length; the chunk's last run gets its length from
// the next chunk that starts a run (pass 3, a thread per chunk).
// ------------------------------------------------------------------
constexpr uint64_t RLE_CCH = 16384;

__global__ __launch_bounds__(64) void k_rle_count(const uint64_t *v, const uint64_t *d_m, uint64_t nch_bound,
                                                  uint64_t *cnt, uint64_t *first)
{
    const uint64_t m = *d_m;
    const uint64_t nch = (m + RLE_CCH - 1) / RLE_CCH;
    const uint32_t t = threadIdx.x;
    for (uint64_t c = blockIdx.x; c < nch_bound; c += gridDim.x) {
        uint64_t runs = 0, f = ~0ull;
        if (c < nch) {
            const uint64_t lo = c * RLE_CCH, hi = lo + RLE_CCH < m ? lo + RLE_CCH : m;
            for (uint64_t i0 = lo; i0 < hi; i0 += 64) {
                const uint64_t i = i0 + t;
                const bool st = i < hi && (i == 0 || v[i] != v[i - 1]);
                const uint64_t b = __ballot(st);
                if (b && f == ~0ull) f = i0 + __ffsll((unsigned long long)b) - 1;
                runs += __popcll(b);
            }
        }
        if (t == 0) {
            cnt[c] = runs;
            first[c] = f;
        }
    }
}

__global__ __launch_bounds__(64) void k_rle_write(const uint64_t *v, const uint64_t *d_m, const uint64_t *roff,
                                                  uint64_t *out_w)
{
    const uint64_t m = *d_m;
    const uint64_t nch = (m + RLE_CCH - 1) / RLE_CCH;
    const uint32_t t = threadIdx.x;
    for (uint64_t c = blockIdx.x; c < nch; c += gridDim.x) {
        const uint64_t lo = c * RLE_CCH, hi = lo + RLE_CCH < m ? lo + RLE_CCH : m;
        uint64_t r0 = roff[c], last = 0;
        bool any = false;
        for (uint64_t i0 = lo; i0 < hi; i0 += 64) {
            const uint64_t i = i0 + t;
            const bool st = i < hi && (i == 0 || v[i] != v[i - 1]);
            const uint64_t b = __ballot(st);
            if (st) {
                const uint64_t below = b & ((1ull << t) - 1ull);
                const uint64_t r = r0 + __popcll(below);
                out_w[2 * r] = v[i];
                // the run before this start: in this chunk, its length
                if (below) out_w[2 * (r - 1) + 1] = i - (i0 + 63 - __clzll(below));
                else if (any) out_w[2 * (r - 1) + 1] = i - last;
            }
            if (b) {
                last = i0 + 63 - __clzll(b);
                any = true;
            }
            r0 += __popcll(b);
        }
        // the chunk's last run: its start, until pass 3 knows the next one
        if (t == 0 && any) out_w[2 * (r0 - 1) + 1] = last;
    }
}

__global__ __launch_bounds__(WG) void k_rle_close(const uint64_t *d_m, const uint64_t *cnt, const uint64_t *first,
                                                  const uint64_t *roff, uint64_t *out_w, ms_dev_meta *meta)
{
    const uint64_t m = *d_m;
    const uint64_t nch = (m + RLE_CCH - 1) / RLE_CCH;
    for (uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x; c < nch; c += (uint64_t)gridDim.x * WG) {
        if (!cnt[c]) continue;
        uint64_t next = m;
        for (uint64_t d = c + 1; d < nch; d++)
            if (cnt[d]) { next = first[d]; break; }
        const uint64_t r = roff[c] + cnt[c] - 1;
        out_w[2 * r + 1] = next - out_w[2 * r + 1];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint64_t runs = nch ? roff[nch - 1] + cnt[nch - 1] : 0;
        meta->count = m;
        meta->size_used = meta->size_compr = 16 * runs;
    }
}

int launch_rle_compress(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, temps &sc)
{
    hipStream_t s = sc.s;
    const uint64_t nch_bound = m_bound / RLE_CCH + 2;
    uint64_t *cnt = (uint64_t *)sc.get(nch_bound * 8 + 64);
    uint64_t *first = (uint64_t *)sc.get(nch_bound * 8 + 64);
    uint64_t *roff = (uint64_t *)sc.get(nch_bound * 8 + 64);
    if (!cnt || !first || !roff) return ms_fail(MS_ERR_NOMEM, "rle scratch");
    const uint32_t grid = grid_for(nch_bound, 1, 65536);
    hipLaunchKernelGGL(k_rle_count, dim3(grid), dim3(64), 0, s, d_v, d_m, nch_bound, cnt, first);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan(cnt, roff, nch_bound, sc));
    hipLaunchKernelGGL(k_rle_write, dim3(grid), dim3(64), 0, s, d_v, d_m, (const uint64_t *)roff,
                       (uint64_t *)out->data);
    hipLaunchKernelGGL(k_rle_close, dim3(grid_for(nch_bound, WG)), dim3(WG), 0, s, d_m, (const uint64_t *)cnt,
                       (const uint64_t *)first, (const uint64_t *)roff, (uint64_t *)out->data, out->meta);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

}  // namespace ms
