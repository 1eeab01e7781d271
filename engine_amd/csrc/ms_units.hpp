// ms_units.hpp -- per-unit decode routines shared by the kernels of one
// column (ms_decode.hip) and the batched persistent kernel (ms_batch.hip).
// A unit is 512 values decoded by one wave (see ms_device.hpp for the
// thread mapping of the vertical bit-packed layouts).
#pragma once

#include <type_traits>

#include "ms_kernels.h"

namespace ms {

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

constexpr int WG = 256;
constexpr int WAVES = WG / MS_WAVE;

inline uint32_t grid_for(uint64_t items, uint64_t per_wg, uint32_t cap = 8192)
{
    uint64_t g = (items + per_wg - 1) / per_wg;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

__device__ __forceinline__ uint32_t wave_index() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ __forceinline__ uint64_t uniform64(uint64_t x)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ ms_parts parts_of(const ms_colv &c)
{
    return ms_parts_of(c.g, c.meta->count, c.meta->size_used, c.meta->size_compr);
}

// Writes the workgroup's sum of per-thread values to part[blockIdx.x]
// (no same-address atomics: a finishing kernel adds the partial sums).
__device__ __forceinline__ void wg_partial(uint64_t acc, uint64_t *part)
{
    __shared__ uint64_t s_part[WAVES];
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t s = 0;
#pragma unroll
        for (int w = 0; w < WAVES; w++) s += s_part[w];
        part[blockIdx.x] = s;
    }
}

// ------------------------------------------------------------------
// Random read of value i of the compressed part (static_vbp, or any
// indexed layout; for_f adds the block's reference value).
// ------------------------------------------------------------------
__device__ __forceinline__ uint64_t read_compr(const ms_colv &c, uint64_t i)
{
    if (c.g.kind == MS_STATIC_VBP) {
        const uint32_t S = c.g.lanes;
        return vbp_read((const uint64_t *)c.data, S, (uint32_t)(i % S), i / S, c.g.bw);
    }
    const uint64_t b = i / c.g.inner_bs;
    const uint64_t r = i - b * c.g.inner_bs;
    const uint64_t e = gload(c.index + b);
    const uint32_t bw = (uint32_t)(e & 0xff);
    const uint32_t S = c.g.lanes;
    uint64_t v = vbp_read((const uint64_t *)(c.data + (e >> 8)), S, (uint32_t)(r % S), r / S, bw);
    if (c.g.kind == MS_FOR) v += gload(c.refs + b / (c.g.outer_block / c.g.inner_bs));
    return v;
}

// ------------------------------------------------------------------
// Fused select + sum over 8 codes of one thread.
// ------------------------------------------------------------------
// Select + sum of the 8 codes cut from preloaded words (BW <= 32).
template <int BW>
__device__ __forceinline__ void ss_words(const uint64_t *w, uint32_t sh, const ms_cpred &cp, uint64_t ref,
                                         uint64_t &acc)
{
    using A = typename std::conditional<(BW <= 29), uint32_t, uint64_t>::type;
    uint32_t c[8];
    vbp_cut32<BW>(w, sh, c);
    A s = 0;
    uint32_t cnt;
    if (cp.mode == PM_ALL) {
#pragma unroll
        for (int m = 0; m < 8; m++) s += c[m];
        cnt = 8;
    } else {
        const uint32_t lo = (uint32_t)cp.clo, span = (uint32_t)cp.cspan;
        const bool out = cp.mode == PM_OUT;
        cnt = 0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const bool take = ((uint32_t)(c[m] - lo) <= span) != out;
            s += take ? c[m] : 0u;
            cnt += take;
        }
    }
    acc += (uint64_t)s + (uint64_t)cnt * ref;
}

// Select + sum of the 8 codes in a thread's loaded words (any BW).
template <int BW>
__device__ __forceinline__ void ss_loaded(const uint64_t (&w)[vbp_traits<BW>::NW], uint32_t sh, const ms_cpred &cp,
                                          uint64_t ref, uint64_t &acc)
{
    if constexpr (BW <= 32) {
        ss_words<BW>(w, sh, cp, ref, acc);
    } else {
        uint64_t c[8];
        vbp_cut64<BW>(w, sh, c);
        uint64_t s = 0, cnt = 0;
        if (cp.mode == PM_ALL) {
#pragma unroll
            for (int m = 0; m < 8; m++) s += c[m];
            cnt = 8;
        } else {
            const bool out = cp.mode == PM_OUT;
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const bool take = ((c[m] - cp.clo) <= cp.cspan) != out;
                s += take ? c[m] : 0ull;
                cnt += take;
            }
        }
        acc += s + cnt * ref;
    }
}

template <int BW>
__device__ __forceinline__ void ss_unit(const uint64_t *base, uint32_t S, uint32_t j, uint64_t g, const ms_cpred &cp,
                                        uint64_t ref, uint64_t &acc)
{
    uint64_t w[vbp_traits<BW>::NW];
    uint32_t sh;
    vbp_load<BW>(base, S, j, g, w, sh);
    ss_loaded<BW>(w, sh, cp, ref, acc);
}

#define MS_BW_CASES(X)                                                                                              \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)     \
    X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39)  \
    X(40) X(41) X(42) X(43) X(44) X(45) X(46) X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58)  \
    X(59) X(60) X(61) X(62) X(63) X(64)

__device__ __forceinline__ void ss_dispatch(uint32_t bw, const uint64_t *base, uint32_t S, uint32_t j, uint64_t g,
                                            const ms_cpred &cp, uint64_t ref, uint64_t &acc)
{
    switch (bw) {
#define MS_CASE(B) \
    case B: ss_unit<B>(base, S, j, g, cp, ref, acc); break;
        MS_BW_CASES(MS_CASE)
#undef MS_CASE
    default: break;
    }
}

#define MS_BW32_CASES(X)                                                                                            \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)     \
    X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

__device__ __forceinline__ void ss_words_dispatch(uint32_t bw, const uint64_t *w, uint32_t sh, const ms_cpred &cp,
                                                  uint64_t ref, uint64_t &acc)
{
    switch (bw) {
#define MS_CASE(B) \
    case B: ss_words<B>(w, sh, cp, ref, acc); break;
        MS_BW32_CASES(MS_CASE)
#undef MS_CASE
    default: break;
    }
}

// Decodes 8 codes as 64-bit values with a run-time bit width.
template <int BW>
__device__ __forceinline__ void dec_unit(const uint64_t *base, uint32_t S, uint32_t j, uint64_t g, uint64_t (&c)[8])
{
    uint64_t w[vbp_traits<BW>::NW];
    uint32_t sh;
    vbp_load<BW>(base, S, j, g, w, sh);
    vbp_cut64<BW>(w, sh, c);
}

__device__ __forceinline__ void dec_dispatch(uint32_t bw, const uint64_t *base, uint32_t S, uint32_t j, uint64_t g,
                                             uint64_t (&c)[8])
{
    switch (bw) {
#define MS_CASE(B) \
    case B: dec_unit<B>(base, S, j, g, c); break;
        MS_BW_CASES(MS_CASE)
#undef MS_CASE
    default:
#pragma unroll
        for (int m = 0; m < 8; m++) c[m] = 0;
        break;
    }
}

// Indexed layouts with 8 lanes and blocks of a multiple of 512 values:
// dynamic_vbp_f<512k, P, 8> and for_f<L, P, dynamic_vbp_f<512k, P', 8>>.
// One wave decodes a 512-value unit; the unit's index entry for the next
// iteration is fetched one iteration ahead.  Per unit the predicate is
// specialised on the block's code range: no value / every value / x <= c /
// x in [lo, hi] / x outside [lo, hi].
enum { SM_NONE = 0, SM_ALL = 1, SM_LE = 2, SM_IN = 3, SM_OUT = 4 };

// Words of lane j holding its codes 8g .. 8g+7 (BW <= 32: at most 5
// words).  The last word may lie past the thread's bits (and past the
// block): column buffers are padded, and those bits are never used.
template <int BW>
__device__ __forceinline__ void unit_words(const uint64_t *__restrict__ base, uint32_t j, uint32_t g,
                                           uint64_t (&w)[5])
{
    constexpr int NW = vbp_traits<BW>::NW;
    static_assert(BW <= 32 && NW <= 5, "at most 5 words for 32-bit codes");
    const uint32_t byte0 = g * BW;
    const uint64_t *p = base + (uint64_t)(byte0 >> 3) * 8 + j;
#pragma unroll
    for (int q = 0; q < NW; q++) w[q] = gload(p + q * 8);
}

__device__ __forceinline__ void unit_words_dispatch(uint32_t bw, const uint64_t *base, uint32_t j, uint32_t g,
                                                    uint64_t (&w)[5])
{
    switch (bw) {
#define MS_CASE(B) \
    case B: unit_words<B>(base, j, g, w); break;
        MS_BW32_CASES(MS_CASE)
#undef MS_CASE
    default: break;
    }
}

// Select + sum of the thread's 8 codes cut from its words, under the
// unit's predicate mode; CNT adds ref once per selected code (for_f).
template <int BW, bool CNT>
__device__ __forceinline__ void ss_cut_sum(const uint64_t (&w)[5], uint32_t g, uint32_t mode, uint32_t lo,
                                           uint32_t span, uint64_t ref, uint64_t &acc)
{
    uint32_t c[8];
    vbp_cut32<BW>(w, ((g * BW) & 7) * 8, c);
    using A = typename std::conditional<(BW <= 29), uint32_t, uint64_t>::type;
    A s = 0;
    uint32_t cnt = 0;
    if (mode == SM_ALL) {
#pragma unroll
        for (int m = 0; m < 8; m++) s += c[m];
        cnt = 8;
    } else if (mode == SM_LE) {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const bool take = c[m] <= span;
            s += take ? c[m] : 0u;
            if (CNT) cnt += take;
        }
    } else {
        const bool out = mode == SM_OUT;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const bool take = ((uint32_t)(c[m] - lo) <= span) != out;
            s += take ? c[m] : 0u;
            if (CNT) cnt += take;
        }
    }
    acc += (uint64_t)s;
    if (CNT) acc += (uint64_t)cnt * ref;
}

// Per-bit-width specialisation of a predicate for columns whose values are
// their codes (no reference value): built on the host once per call.
struct ms_bwtab {
    uint32_t mode[33];
    uint32_t lo[33];
    uint32_t span[33];
};

// ms_bwtab for codes that are the values (ref 0), bit widths 1..32.
inline ms_bwtab make_bwtab(const ms_pred &p)
{
    ms_bwtab t = {};
    for (uint32_t bw = 1; bw <= 32; bw++) {
        const uint64_t maxcode = (1ull << bw) - 1ull;
        uint64_t lo = p.lo, hi = p.hi;
        const bool in_empty = p.empty || lo > maxcode;
        if (!in_empty && hi > maxcode) hi = maxcode;
        const bool in_all = !in_empty && lo == 0 && hi == maxcode;
        uint32_t mode;
        if (!p.neg) mode = in_empty ? SM_NONE : in_all ? SM_ALL : lo == 0 ? SM_LE : SM_IN;
        else mode = in_empty ? SM_ALL : in_all ? SM_NONE : SM_OUT;
        t.mode[bw] = mode;
        t.lo[bw] = in_empty ? 0 : (uint32_t)lo;
        t.span[bw] = in_empty ? 0 : (uint32_t)(hi - lo);
    }
    return t;
}

// Predicates of the form x <= hi (LT, LE: not negated, not empty, lo = 0).
__host__ __device__ inline bool pred_le_form(const ms_pred &p) { return !p.neg && !p.empty && p.lo == 0; }

// Per-unit mode of such a predicate for x = ref + code, code in [0, maxcode]
// without wrap-around: a few scalar operations instead of code_pred.
__device__ __forceinline__ uint32_t unit_mode_le(const ms_pred &p, uint64_t ref, uint64_t maxcode, uint32_t &lo,
                                                 uint32_t &span)
{
    lo = 0;
    span = 0;
    if (ref > p.hi) return SM_NONE;
    const uint64_t s = p.hi - ref;
    if (s >= maxcode) return SM_ALL;
    span = (uint32_t)s;
    return SM_LE;
}

template <bool CNT>
__device__ __forceinline__ void ss_cut_sum_dispatch(uint32_t bw, const uint64_t (&w)[5], uint32_t g, uint32_t mode,
                                                    uint32_t lo, uint32_t span, uint64_t ref, uint64_t &acc)
{
    switch (bw) {
#define MS_CASE(B) \
    case B: ss_cut_sum<B, CNT>(w, g, mode, lo, span, ref, acc); break;
        MS_BW32_CASES(MS_CASE)
#undef MS_CASE
    default: break;
    }
}

// Value-by-value select + sum of a thread's 8 codes (wide or wrapping):
// rare, kept out of line so it does not set the kernels' register budget
// (and declared pure, so the callers' uniform loads stay scalar).
__device__ __attribute__((noinline, pure)) uint64_t slow_codes(const ms_pred p, const uint64_t *base, uint32_t j,
                                                               uint64_t g, uint32_t bw, uint64_t ref)
{
    uint64_t acc = 0;
#pragma unroll 1
    for (int m = 0; m < 8; m++) {
        const uint64_t x = ref + vbp_read(base, 8, j, 8 * g + m, bw);
        if (pred_eval(p, x)) acc += x;
    }
    return acc;
}

enum { SM_SLOW = 5 };   // wide (> 32-bit) or wrapping codes: value by value

// Up to four consecutive 512-value units u0 .. u0+cnt-1 of an 8-lane
// indexed layout (dynamic_vbp_f blocks, or for_f's inner blocks): e[u] =
// their block index entries, r[u] = their for_f reference values, mask =
// units per block - 1; thread (gl, j) takes codes 8 gl' .. of lane j with
// gl' = its unit's offset in the block * 8 + gl.  Every unit's predicate
// mode is decided and its words requested before the first unit is
// decoded, so a wave keeps up to four units' loads in flight; a unit no
// value of which can pass is not read.
//
// vbp_unit_prep decides unit u's predicate mode and requests its words;
// vbp_unit_finish decodes them.
template <bool IS_FOR, bool LEFORM>
__device__ __forceinline__ void vbp_unit_prep(const uint8_t *data, uint64_t e, uint64_t r, bool valid, uint64_t u,
                                              uint64_t mask, const ms_pred &p, const ms_bwtab &tab, uint32_t j,
                                              uint32_t gl, uint32_t &mode, uint32_t &lo, uint32_t &span,
                                              uint64_t (&w)[5])
{
    mode = SM_NONE;
    lo = span = 0;
    if (!valid) return;
    const uint32_t bw = (uint32_t)(e & 0xff);
    if (IS_FOR) {
        const uint64_t maxcode = bw >= 64 ? ~0ull : (1ull << bw) - 1ull;
        if (bw > 32 || r + maxcode < r) {
            mode = SM_SLOW;
        } else if (LEFORM) {
            mode = unit_mode_le(p, r, maxcode, lo, span);
        } else {
            const ms_cpred cp = code_pred(p, r, maxcode);
            mode = cp.mode == PM_NONE ? SM_NONE : cp.mode == PM_ALL ? SM_ALL : cp.mode == PM_OUT ? SM_OUT
                 : cp.clo == 0 ? SM_LE : SM_IN;
            lo = (uint32_t)cp.clo;
            span = (uint32_t)cp.cspan;
        }
    } else if (bw > 32) {
        mode = SM_SLOW;
    } else {
        mode = tab.mode[bw];
        lo = tab.lo[bw];
        span = tab.span[bw];
    }
    if (mode != SM_NONE && mode != SM_SLOW) {
        uint32_t g = (uint32_t)(u & mask) * 8 + gl;
        asm volatile("" : "+v"(g));   // keeps per-width offsets of g from being hoisted
        unit_words_dispatch(bw, (const uint64_t *)(data + (e >> 8)), j, g, w);
    }
}

template <bool IS_FOR>
__device__ __forceinline__ void vbp_unit_finish(const uint8_t *data, uint64_t e, uint64_t r, uint64_t u,
                                                uint64_t mask, const ms_pred &p, uint32_t j, uint32_t gl,
                                                uint32_t mode, uint32_t lo, uint32_t span, const uint64_t (&w)[5],
                                                uint64_t &acc)
{
    if (mode == SM_NONE) return;
    const uint32_t bw = (uint32_t)(e & 0xff);
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));
    if (mode == SM_SLOW) acc += slow_codes(p, (const uint64_t *)(data + (e >> 8)), j, g, bw, IS_FOR ? r : 0ull);
    else ss_cut_sum_dispatch<IS_FOR>(bw, w, g, mode, lo, span, r, acc);
}

// NU (2 or 4) units: all their words requested before the first is cut.
template <bool IS_FOR, bool LEFORM, int NU = 4>
__device__ __forceinline__ void vbp_units_ss(const uint8_t *data, const uint64_t (&e)[4], const uint64_t (&r)[4],
                                             uint32_t cnt, uint64_t u0, uint64_t mask, const ms_pred &p,
                                             const ms_bwtab &tab, uint32_t j, uint32_t gl, uint64_t &acc)
{
    static_assert(NU == 2 || NU == 4, "units per group");
    // named word arrays (not one indexed array): they stay in registers
    uint32_t m0, m1, m2 = 0, m3 = 0, l0, l1, l2 = 0, l3 = 0, s0, s1, s2 = 0, s3 = 0;
    uint64_t w0[5], w1[5], w2[5], w3[5];
    vbp_unit_prep<IS_FOR, LEFORM>(data, e[0], r[0], cnt > 0, u0, mask, p, tab, j, gl, m0, l0, s0, w0);
    vbp_unit_prep<IS_FOR, LEFORM>(data, e[1], r[1], cnt > 1, u0 + 1, mask, p, tab, j, gl, m1, l1, s1, w1);
    if (NU == 4) {
        vbp_unit_prep<IS_FOR, LEFORM>(data, e[2], r[2], cnt > 2, u0 + 2, mask, p, tab, j, gl, m2, l2, s2, w2);
        vbp_unit_prep<IS_FOR, LEFORM>(data, e[3], r[3], cnt > 3, u0 + 3, mask, p, tab, j, gl, m3, l3, s3, w3);
    }
    vbp_unit_finish<IS_FOR>(data, e[0], r[0], u0, mask, p, j, gl, m0, l0, s0, w0, acc);
    vbp_unit_finish<IS_FOR>(data, e[1], r[1], u0 + 1, mask, p, j, gl, m1, l1, s1, w1, acc);
    if (NU == 4) {
        vbp_unit_finish<IS_FOR>(data, e[2], r[2], u0 + 2, mask, p, j, gl, m2, l2, s2, w2, acc);
        vbp_unit_finish<IS_FOR>(data, e[3], r[3], u0 + 3, mask, p, j, gl, m3, l3, s3, w3, acc);
    }
}

// ------------------------------------------------------------------
// rle_f select + sum over runs [r0, r1) of the (value, length) pairs, by
// one wave: lane t takes runs r0 + t + 64 i, four loads in flight.
// ------------------------------------------------------------------
__device__ __forceinline__ uint64_t rle_runs_ss(const u64x2 *runs, uint64_t r0, uint64_t r1, const ms_pred &p,
                                                uint32_t t)
{
    uint64_t acc = 0;
    uint64_t k = r0;
    for (; k + 256 <= r1; k += 256) {
        u64x2 v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = gload(runs + k + 64 * q + t);
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (pred_eval(p, v[q].x)) acc += v[q].x * v[q].y;
    }
    for (k += t; k < r1; k += 64) {
        const u64x2 vl = gload(runs + k);
        if (pred_eval(p, vl.x)) acc += vl.x * vl.y;
    }
    return acc;
}

// ------------------------------------------------------------------
// delta_f select + sum, one wave per unit of deltas with the lanes'
// running sums ("carries") at the unit's start.  Thread t handles delta
// lane j = t / 8 and code group gl = t % 8, so each lane's 64 codes sit in
// 8 consecutive threads and its running sum is a scan within 8-thread
// groups done with DPP row shifts (no LDS traffic).  Deltas of up to 26
// bits are decoded and summed in 32-bit arithmetic (64 codes of a lane
// sum to less than 2^32); most lanes' values lie wholly inside or outside
// the predicate's range, and those need no scan (delta_codes_ss).
// ------------------------------------------------------------------

// Value of the lane k (1, 2 or 4) places below within its 8-thread group,
// or 0 (DPP row_shr with bound control, masked at the group's start).
template <int K>
__device__ __forceinline__ uint32_t group8_up(uint32_t x, uint32_t t)
{
    const uint32_t y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + K, 0xf, 0xf, true);
    return (t & 7) >= (uint32_t)K ? y : 0u;
}

// Inclusive scan within each 8-thread group.
__device__ __forceinline__ uint32_t group8_scan(uint32_t x, uint32_t t)
{
    x += group8_up<1>(x, t);
    x += group8_up<2>(x, t);
    x += group8_up<4>(x, t);
    return x;
}

// Sum over each 8-thread group, in every thread of the group (quad swaps,
// then the half-row mirror pairs the two quads).
__device__ __forceinline__ uint32_t group8_sum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xb1, 0xf, 0xf, false);    // quad_perm [1,0,3,2]
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4e, 0xf, 0xf, false);    // quad_perm [2,3,0,1]
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xf, 0xf, false);   // row_half_mirror
    return x;
}

__device__ __forceinline__ uint64_t group8_scan64(uint64_t x, uint32_t t)
{
#pragma unroll
    for (uint32_t o = 1; o < 8; o <<= 1) {
        const uint64_t v = __shfl_up(x, o, MS_WAVE);
        if ((t & 7) >= o) x += v;
    }
    return x;
}
#define MS_BW26_CASES(X)                                                                                            \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20)     \
    X(21) X(22) X(23) X(24) X(25) X(26)

// Cuts the 8 codes (<= 26 bits) out of words loaded by unit_words_dispatch.
__device__ __forceinline__ void dec32_cut_dispatch(uint32_t bw, const uint64_t (&w)[5], uint32_t g, uint32_t (&c)[8])
{
    switch (bw) {
#define MS_CASE(B) \
    case B: vbp_cut32<B>(w, ((g * B) & 7) * 8, c); break;
        MS_BW26_CASES(MS_CASE)
#undef MS_CASE
    default:
#pragma unroll
        for (int m = 0; m < 8; m++) c[m] = 0;
        break;
    }
}

// Deltas of more than 26 bits, in 64-bit arithmetic: rare (unsorted
// data), kept out of line so it does not set the callers' register budget.
// Returns {select + sum of the thread's 8 values, the lane's new carry}.
__device__ __forceinline__ u64x2 delta_unit_wide(const ms_pred p, const uint64_t *base, uint32_t bw,
                                                                  uint32_t j, uint32_t g, uint32_t t, uint64_t carry)
{
    uint64_t d[8];
    dec_dispatch(bw, base, 8, j, g, d);
#pragma unroll
    for (int m = 1; m < 8; m++) d[m] += d[m - 1];
    const uint64_t tot = d[7];
    const uint64_t incl = group8_scan64(tot, t);
    const uint64_t bt = carry + incl - tot;
    uint64_t acc = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const uint64_t x = bt + d[m];
        acc += pred_eval(p, x) ? x : 0ull;
    }
    u64x2 r;
    r.x = acc;
    r.y = carry + __shfl(incl, (t & ~7u) + 7, MS_WAVE);
    return r;
}

// The same, out of line: kernels that decode many layouts keep their
// register budget (and their scalar loads: it is pure).
__device__ __attribute__((noinline, pure)) u64x2 delta_unit_wide_call(const ms_pred p, const uint64_t *base,
                                                                       uint32_t bw, uint32_t j, uint32_t g,
                                                                       uint32_t t, uint64_t carry)
{
    return delta_unit_wide(p, base, bw, j, g, t, carry);
}

// Select + sum of one unit from the thread's 8 deltas (<= 26 bits each);
// carry = lane j's running sum before the unit, advanced past it.  Lane
// j's 64 values in the unit are carry plus running sums of its deltas:
// non-decreasing, inside [carry, carry + lane total] when that does not
// wrap.  A lane whose whole range is inside or outside the predicate's
// interval needs no scan: its 8 threads' values add up to
//   sum over g of (8 carry + 8 (7 - g) tot_g + psum_g)
// (tot_g = thread g's delta total, psum_g = the sum of its 8 running
// sums), each thread adding its own term.  Only a wave holding a lane that
// straddles a bound (rare) scans and tests per thread, then per value.
__device__ __forceinline__ void delta_codes_ss(const ms_pred &p, uint32_t (&d)[8], uint32_t gl, uint32_t t,
                                               uint64_t &carry, uint64_t &acc)
{
    uint32_t psum = d[0];
#pragma unroll
    for (int m = 1; m < 8; m++) {
        d[m] += d[m - 1];
        psum += d[m];
    }
    const uint32_t tot = d[7];
    const uint64_t c0 = carry, c1 = carry + group8_sum(tot);
    const uint64_t span = p.hi - p.lo;
    bool whole, sel;
    if (pred_le_form(p)) {   // x <= hi (uniform): two compares
        const bool in0 = c0 <= p.hi, in1 = c1 <= p.hi;
        whole = c1 >= c0 && in0 == in1;
        sel = in0;
    } else {
        const bool in0 = !p.empty && c0 - p.lo <= span;
        const bool in1 = !p.empty && c1 - p.lo <= span;
        whole = c1 >= c0 && in0 == in1 && (p.empty || !(c0 < p.lo && c1 > p.hi));
        sel = in0 != (bool)p.neg;
    }
    if (__builtin_amdgcn_ballot_w64(!whole) == 0) {
        if (sel) acc += 8ull * c0 + (uint64_t)(8u * (7u - gl)) * tot + psum;
    } else {
        const uint32_t incl = group8_scan(tot, t);
        if (whole) {
            if (sel) acc += 8ull * c0 + (uint64_t)(8u * (7u - gl)) * tot + psum;
        } else {
            // the thread's values lie in [bt, be]
            const uint64_t bt = c0 + (incl - tot), be = bt + tot;
            const bool i0 = !p.empty && bt - p.lo <= span;
            const bool i1 = !p.empty && be - p.lo <= span;
            if (be >= bt && i0 == i1 && (p.empty || !(bt < p.lo && be > p.hi))) {
                if (i0 != (bool)p.neg) acc += 8ull * bt + psum;
            } else {
#pragma unroll 1
                for (int m = 0; m < 8; m++) {
                    const uint64_t x = bt + d[m];
                    acc += pred_eval(p, x) ? x : 0ull;
                }
            }
        }
    }
    carry = c1;
}

// Requests the words of delta unit u (entry e) and decodes them (below).
__device__ __forceinline__ void delta_unit_prep(const uint8_t *data, uint64_t e, bool valid, uint64_t u,
                                                uint64_t mask, uint32_t j, uint32_t gl, uint64_t (&w)[5])
{
    const uint32_t bw = (uint32_t)(e & 0xff);
    if (!valid || bw > 26) return;
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));   // keeps per-width offsets of g from being hoisted
    unit_words_dispatch(bw, (const uint64_t *)(data + (e >> 8)), j, g, w);
}

template <bool WIDE_OUT_OF_LINE>
__device__ __forceinline__ void delta_unit_finish(const uint8_t *data, uint64_t e, bool valid, uint64_t u,
                                                  uint64_t mask, const ms_pred &p, uint32_t t, uint32_t j,
                                                  uint32_t gl, const uint64_t (&w)[5], uint64_t &carry,
                                                  uint64_t &acc)
{
    if (!valid) return;
    const uint32_t bw = (uint32_t)(e & 0xff);
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));
    if (bw <= 26) {
        uint32_t d[8];
        dec32_cut_dispatch(bw, w, g, d);
        delta_codes_ss(p, d, gl, t, carry, acc);
    } else {
        const uint64_t *base = (const uint64_t *)(data + (e >> 8));
        const u64x2 r = WIDE_OUT_OF_LINE ? delta_unit_wide_call(p, base, bw, j, g, t, carry)
                                         : delta_unit_wide(p, base, bw, j, g, t, carry);
        acc += r.x;
        carry = r.y;
    }
}

// Units u0 .. u0+cnt-1 (cnt <= 4: one delta tile) of an 8-lane delta_f
// column with 512-value-multiple inner blocks; e[u] = their block index
// entries, mask = units per inner block - 1, carry = lane j's running sum
// at u0.  The words of all units are requested before the first unit is
// decoded, so a wave keeps four units' loads in flight.
template <bool WIDE_OUT_OF_LINE>
__device__ __forceinline__ void delta_units_ss(const uint8_t *data, const uint64_t (&e)[4], uint32_t cnt,
                                               uint64_t u0, uint64_t mask, const ms_pred &p, uint32_t t,
                                               uint32_t j, uint32_t gl, uint64_t &carry, uint64_t &acc)
{
    uint64_t w0[5], w1[5], w2[5], w3[5];   // named arrays: they stay in registers
    delta_unit_prep(data, e[0], cnt > 0, u0, mask, j, gl, w0);
    delta_unit_prep(data, e[1], cnt > 1, u0 + 1, mask, j, gl, w1);
    delta_unit_prep(data, e[2], cnt > 2, u0 + 2, mask, j, gl, w2);
    delta_unit_prep(data, e[3], cnt > 3, u0 + 3, mask, j, gl, w3);
    delta_unit_finish<WIDE_OUT_OF_LINE>(data, e[0], cnt > 0, u0, mask, p, t, j, gl, w0, carry, acc);
    delta_unit_finish<WIDE_OUT_OF_LINE>(data, e[1], cnt > 1, u0 + 1, mask, p, t, j, gl, w1, carry, acc);
    delta_unit_finish<WIDE_OUT_OF_LINE>(data, e[2], cnt > 2, u0 + 2, mask, p, t, j, gl, w2, carry, acc);
    delta_unit_finish<WIDE_OUT_OF_LINE>(data, e[3], cnt > 3, u0 + 3, mask, p, t, j, gl, w3, carry, acc);
}

// Block index entries of a delta tile's units, in lanes 0..3 (0 past
// `nunits`).
__device__ __forceinline__ uint64_t delta_tile_entries(const uint64_t *index, uint32_t lg_spb, uint64_t tile,
                                                       uint64_t nunits, uint32_t t)
{
    const uint64_t u = tile * 4 + t;
    return t < 4 && u < nunits ? gload(index + (u >> lg_spb)) : 0ull;
}

// A delta tile (MS_DELTA_TILE = 4 units) of a column with 512-value-
// multiple inner blocks (data, tile carries, log2 units per block), its
// units' index entries in lanes 0..3 of my_e; units at or past `nunits`
// are skipped.
template <bool WIDE_OUT_OF_LINE>
__device__ __forceinline__ void delta_tile_ss(const uint8_t *data, uint64_t my_e, const uint64_t *refs,
                                              uint32_t lg_spb, const ms_pred &p, uint64_t tile, uint64_t nunits,
                                              uint32_t t, uint32_t j, uint32_t gl, uint64_t &acc)
{
    uint64_t carry = gload(refs + tile * 8 + j);
    const uint64_t u0 = tile * 4;
    const uint32_t cnt = nunits - u0 < 4 ? (uint32_t)(nunits - u0) : 4u;
    uint64_t e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) e[u] = readlane64(my_e, u);
    delta_units_ss<WIDE_OUT_OF_LINE>(data, e, cnt, u0, (1ull << lg_spb) - 1ull, p, t, j, gl, carry, acc);
}

// ------------------------------------------------------------------
// Decompression of 512-value units of an 8-lane indexed layout (and of
// delta tiles) into an uncompressed array.  As for select + sum, the words
// of up to four units are requested before the first is decoded, so a wave
// keeps four units' loads in flight; thread (gl, j) writes its 8 values to
// out[u * 512 + 64 gl + 8 m + j].
// ------------------------------------------------------------------
template <bool NT>
__device__ __forceinline__ void st64(uint64_t *p, uint64_t v)
{
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Stores of a thread's 8 values v[m] to o[8 m]: with STORE16 the values
// of two neighbouring lanes j, j + 1 (in threads t and t ^ XOR) are paired
// through a DPP swap, so each thread writes 4 aligned 16-byte pairs (the
// even lane the even m, the odd lane the odd m) and a wave's store
// instruction covers whole 128-byte lines.
#ifndef MS_DECOMP_STORE16
#define MS_DECOMP_STORE16 1
#endif

template <int XOR>
__device__ __forceinline__ uint64_t swap_pair(uint64_t x)
{
    // XOR 1: quad_perm [1,0,3,2]; XOR 8: row_ror:8 (halves of a 16-lane row)
    constexpr int ctl = XOR == 1 ? 0xb1 : 0x128;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, ctl, 0xf, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), ctl, 0xf, 0xf, false);
    return ((uint64_t)hi << 32) | lo;
}

// o points at lane j's slot of row 0 (o = base + 64 gl + j); odd = j & 1.
template <bool NT, int XOR>
__device__ __forceinline__ void store8(uint64_t *o, const uint64_t (&v)[8], uint32_t odd)
{
    if (!MS_DECOMP_STORE16) {
#pragma unroll
        for (int m = 0; m < 8; m++) st64<NT>(o + 8 * m, v[m]);
        return;
    }
    u64x2 *p = (u64x2 *)(o - odd);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t r = swap_pair<XOR>(odd ? v[2 * k] : v[2 * k + 1]);
        u64x2 pr;
        pr.x = odd ? r : v[2 * k];
        pr.y = odd ? v[2 * k + 1] : r;
        u64x2 *q = (u64x2 *)((uint64_t *)p + 8 * (2 * k + odd));
        if (NT) __builtin_nontemporal_store(pr, q);
        else *q = pr;
    }
}

__device__ __forceinline__ void cut32_dispatch(uint32_t bw, const uint64_t (&w)[5], uint32_t g, uint32_t (&c)[8])
{
    switch (bw) {
#define MS_CASE(B) \
    case B: vbp_cut32<B>(w, ((g * B) & 7) * 8, c); break;
        MS_BW32_CASES(MS_CASE)
#undef MS_CASE
    default:
#pragma unroll
        for (int m = 0; m < 8; m++) c[m] = 0;
        break;
    }
}

// Codes of more than 32 bits: rare, out of line (keeps the callers'
// register budget).
template <bool NT>
__device__ __attribute__((noinline)) void dec_wide_store(const uint64_t *base, uint32_t bw, uint32_t j, uint32_t g,
                                                         uint64_t ref, uint64_t *o)
{
    uint64_t d[8];
    dec_dispatch(bw, base, 8, j, g, d);
#pragma unroll
    for (int m = 0; m < 8; m++) d[m] += ref;
    store8<NT, 1>(o, d, j & 1);
}

__device__ __forceinline__ void dec_unit_prep(const uint8_t *data, uint64_t e, bool valid, uint64_t u, uint64_t mask,
                                              uint32_t j, uint32_t gl, uint64_t (&w)[5])
{
    const uint32_t bw = (uint32_t)(e & 0xff);
    if (!valid || bw > 32) return;
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));   // keeps per-width offsets of g from being hoisted
    unit_words_dispatch(bw, (const uint64_t *)(data + (e >> 8)), j, g, w);
}

template <bool NT>
__device__ __forceinline__ void dec_unit_finish(const uint8_t *data, uint64_t e, uint64_t ref, bool valid, uint64_t u,
                                                uint64_t mask, uint32_t j, uint32_t gl, const uint64_t (&w)[5],
                                                uint64_t *out)
{
    if (!valid) return;
    const uint32_t bw = (uint32_t)(e & 0xff);
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));
    uint64_t *o = out + u * 512 + 64 * gl + j;
    if (bw <= 32) {
        uint32_t c[8];
        cut32_dispatch(bw, w, g, c);
        uint64_t v[8];
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = (uint64_t)c[m] + ref;
        store8<NT, 1>(o, v, j & 1);
    } else {
        dec_wide_store<NT>((const uint64_t *)(data + (e >> 8)), bw, j, g, ref, o);
    }
}

// Units u0 .. u0+cnt-1 (cnt <= 4); e[u] = their index entries, r[u] their
// reference values (0 but for for_f).
template <bool NT>
__device__ __forceinline__ void dec_units(const uint8_t *data, const uint64_t (&e)[4], const uint64_t (&r)[4],
                                          uint32_t cnt, uint64_t u0, uint64_t mask, uint32_t j, uint32_t gl,
                                          uint64_t *out)
{
    uint64_t w0[5], w1[5], w2[5], w3[5];   // named arrays: they stay in registers
    dec_unit_prep(data, e[0], cnt > 0, u0, mask, j, gl, w0);
    dec_unit_prep(data, e[1], cnt > 1, u0 + 1, mask, j, gl, w1);
    dec_unit_prep(data, e[2], cnt > 2, u0 + 2, mask, j, gl, w2);
    dec_unit_prep(data, e[3], cnt > 3, u0 + 3, mask, j, gl, w3);
    dec_unit_finish<NT>(data, e[0], r[0], cnt > 0, u0, mask, j, gl, w0, out);
    dec_unit_finish<NT>(data, e[1], r[1], cnt > 1, u0 + 1, mask, j, gl, w1, out);
    dec_unit_finish<NT>(data, e[2], r[2], cnt > 2, u0 + 2, mask, j, gl, w2, out);
    dec_unit_finish<NT>(data, e[3], r[3], cnt > 3, u0 + 3, mask, j, gl, w3, out);
}

// delta_f: thread (j = t / 8, gl = t % 8) holds 8 deltas (<= 26 bits) of
// lane j; writes lane j's running sums from `carry` and advances it.
template <bool NT>
__device__ __forceinline__ void delta_codes_store(uint32_t (&d)[8], uint32_t t, uint64_t &carry, uint64_t *o)
{
#pragma unroll
    for (int m = 1; m < 8; m++) d[m] += d[m - 1];
    const uint32_t tot = d[7];
    const uint32_t incl = group8_scan(tot, t);
    const uint64_t bt = carry + (incl - tot);
    uint64_t v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = bt + d[m];
    store8<NT, 8>(o, v, (t >> 3) & 1);
    carry += group8_sum(tot);
}

template <bool NT>
__device__ __attribute__((noinline)) uint64_t delta_wide_store(const uint64_t *base, uint32_t bw, uint32_t j,
                                                               uint32_t g, uint32_t t, uint64_t carry, uint64_t *o)
{
    uint64_t d[8];
    dec_dispatch(bw, base, 8, j, g, d);
#pragma unroll
    for (int m = 1; m < 8; m++) d[m] += d[m - 1];
    const uint64_t tot = d[7];
    const uint64_t incl = group8_scan64(tot, t);
    const uint64_t bt = carry + incl - tot;
#pragma unroll
    for (int m = 0; m < 8; m++) d[m] += bt;
    store8<NT, 8>(o, d, (t >> 3) & 1);
    return carry + __shfl(incl, (int)((t & ~7u) + 7), MS_WAVE);
}

template <bool NT>
__device__ __forceinline__ void delta_unit_store(const uint8_t *data, uint64_t e, bool valid, uint64_t u,
                                                 uint64_t mask, uint32_t t, uint32_t j, uint32_t gl,
                                                 const uint64_t (&w)[5], uint64_t &carry, uint64_t *out)
{
    if (!valid) return;
    const uint32_t bw = (uint32_t)(e & 0xff);
    uint32_t g = (uint32_t)(u & mask) * 8 + gl;
    asm volatile("" : "+v"(g));
    uint64_t *o = out + u * 512 + 64 * gl + j;
    if (bw <= 26) {
        uint32_t d[8];
        dec32_cut_dispatch(bw, w, g, d);
        delta_codes_store<NT>(d, t, carry, o);
    } else {
        carry = delta_wide_store<NT>((const uint64_t *)(data + (e >> 8)), bw, j, g, t, carry, o);
    }
}

// A delta tile (4 units) into out: entries in lanes 0..3 of my_e.
template <bool NT>
__device__ __forceinline__ void delta_tile_store(const uint8_t *data, uint64_t my_e, const uint64_t *refs,
                                                 uint32_t lg_spb, uint64_t tile, uint64_t nunits, uint32_t t,
                                                 uint32_t j, uint32_t gl, uint64_t *out)
{
    uint64_t carry = gload(refs + tile * 8 + j);
    const uint64_t u0 = tile * 4;
    const uint32_t cnt = nunits - u0 < 4 ? (uint32_t)(nunits - u0) : 4u;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    uint64_t e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) e[u] = readlane64(my_e, u);
    uint64_t w0[5], w1[5], w2[5], w3[5];
    delta_unit_prep(data, e[0], cnt > 0, u0, mask, j, gl, w0);
    delta_unit_prep(data, e[1], cnt > 1, u0 + 1, mask, j, gl, w1);
    delta_unit_prep(data, e[2], cnt > 2, u0 + 2, mask, j, gl, w2);
    delta_unit_prep(data, e[3], cnt > 3, u0 + 3, mask, j, gl, w3);
    delta_unit_store<NT>(data, e[0], cnt > 0, u0, mask, t, j, gl, w0, carry, out);
    delta_unit_store<NT>(data, e[1], cnt > 1, u0 + 1, mask, t, j, gl, w1, carry, out);
    delta_unit_store<NT>(data, e[2], cnt > 2, u0 + 2, mask, t, j, gl, w2, carry, out);
    delta_unit_store<NT>(data, e[3], cnt > 3, u0 + 3, mask, t, j, gl, w3, carry, out);
}

}  // namespace ms
