// ms_device.hpp -- device primitives for gfx950 (wave64).
//
// Vertical bit packing vbp_l<BW, S> (vbp_format_rndaccess.h:60-87): value i
// is code k = i / S of lane j = i % S; lane j's 64-bit word w sits at word
// w*S + j; code k occupies bits [k*BW, (k+1)*BW) of its lane's LSB-first
// bit stream.  A thread decodes 8 consecutive codes k = 8g .. 8g+7 of one
// lane: exactly BW bytes of that lane's stream, starting at byte g*BW.  The
// bytes are loaded as 64-bit words, funnel-shifted to a bit string that
// starts at bit 0, and the 8 codes are cut out at compile-time positions.
#pragma once

#include "ms_internal.h"

#define MS_WAVE 64
#define MS_GLOBAL __attribute__((address_space(1)))

namespace ms {

// Loads through a global-address-space pointer.  Data pointers read from
// memory (work-queue items) are flat pointers to the compiler, and flat
// loads also count on the scalar-load counter, so every scalar-load wait
// would wait for them as well; all column data lives in global memory.
template <class T>
__device__ __forceinline__ T gload(const T *p)
{
    return *(const MS_GLOBAL T *)p;
}

template <int BW>
struct vbp_traits {
    // words a thread may touch: the bit string starts at a byte offset
    // (0..7 bytes into its first word) and is 8*BW bits long.
    static constexpr int NW = (BW % 8 == 0) ? BW / 8 : (8 * BW + 56 + 63) / 64;
    static constexpr int ND = (BW + 3) / 4;        // dwords of the bit string
    static constexpr int NL = (BW + 7) / 8;        // 64-bit words of the bit string
};

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

// Low 32 bits of ((hi:lo) >> s) for s in [0, 32) (one v_alignbit_b32).
__device__ __forceinline__ uint32_t funnel32(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
}

// Loads the words holding codes 8g..8g+7 of lane j.  Words beyond the
// thread's bit string are not read (they may lie past the stream's end).
template <int BW>
__device__ __forceinline__ void vbp_load(const uint64_t *__restrict__ base, uint32_t S, uint32_t j,
                                         uint64_t g, uint64_t (&w)[vbp_traits<BW>::NW], uint32_t &sh)
{
    constexpr int NW = vbp_traits<BW>::NW;
    const uint64_t byte0 = g * (uint64_t)BW;
    const uint64_t w0 = byte0 >> 3;
    sh = (uint32_t)(byte0 & 7) * 8;
    const uint32_t need = sh + 8 * BW;
#pragma unroll
    for (int q = 0; q < NW; q++) {
        if (BW % 8 == 0 || (uint32_t)(q * 64) < need) w[q] = gload(base + (w0 + q) * S + j);
        else w[q] = 0;
    }
}

// Cuts the 8 codes (BW <= 32) out of the loaded words.
template <int BW>
__device__ __forceinline__ void vbp_cut32(const uint64_t *w, uint32_t sh, uint32_t (&c)[8])
{
    static_assert(BW <= 32, "32-bit cut");
    constexpr int NW = vbp_traits<BW>::NW;
    constexpr int ND = vbp_traits<BW>::ND;
    constexpr uint32_t MASK = BW == 32 ? 0xffffffffu : ((1u << BW) - 1u);
    uint32_t D[2 * NW + 2];
#pragma unroll
    for (int q = 0; q < NW; q++) { D[2 * q] = lo32(w[q]); D[2 * q + 1] = hi32(w[q]); }
    D[2 * NW] = 0;
    D[2 * NW + 1] = 0;
    uint32_t L[ND + 1];
    if (BW % 8 == 0) {
#pragma unroll
        for (int i = 0; i < ND; i++) L[i] = D[i];
    } else {
        const bool d0 = sh >= 32;
        const uint32_t s = sh & 31;
#pragma unroll
        for (int i = 0; i < ND; i++) {
            const uint32_t lo = d0 ? D[i + 1] : D[i];
            const uint32_t hi = d0 ? D[i + 2] : D[i + 1];
            L[i] = funnel32(hi, lo, s);
        }
    }
    L[ND] = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int p = m * BW;
        const int a = p >> 5, b = p & 31;
        uint32_t v;
        if (b + BW <= 32) v = L[a] >> b;
        else v = funnel32(L[a + 1], L[a], b);
        c[m] = v & MASK;
    }
}

// Cuts the 8 codes for any BW (1..64) as 64-bit values.
template <int BW>
__device__ __forceinline__ void vbp_cut64(const uint64_t *w, uint32_t sh, uint64_t (&c)[8])
{
    constexpr int NW = vbp_traits<BW>::NW;
    constexpr int NL = vbp_traits<BW>::NL;
    constexpr uint64_t MASK = BW == 64 ? ~0ull : ((1ull << BW) - 1ull);
    uint64_t L[NL + 1];
    if (BW % 8 == 0) {
#pragma unroll
        for (int i = 0; i < NL; i++) L[i] = w[i];
    } else {
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const uint64_t nx = (i + 1 < NW) ? w[i + 1] : 0ull;
            L[i] = sh ? ((w[i] >> sh) | (nx << (64 - sh))) : w[i];
        }
    }
    L[NL] = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int p = m * BW;
        const int a = p >> 6, b = p & 63;
        uint64_t v = L[a] >> b;
        if (b + BW > 64) v |= L[a + 1] << (64 - b);
        c[m] = v & MASK;
    }
}

// Loads, for a run-time bit width bw <= 32, the (at most 5) words holding
// codes 8g..8g+7 of lane j; words past the thread's bit string stay 0.
__device__ __forceinline__ void vbp_load_rt(const uint64_t *__restrict__ base, uint32_t S, uint32_t j, uint64_t g,
                                            uint32_t bw, uint64_t (&w)[5], uint32_t &sh)
{
    const uint64_t byte0 = g * (uint64_t)bw;
    const uint64_t w0 = byte0 >> 3;
    sh = (uint32_t)(byte0 & 7) * 8;
    const uint32_t need = sh + 8 * bw;
#pragma unroll
    for (int q = 0; q < 5; q++)
        w[q] = (uint32_t)(q * 64) < need ? gload(base + (w0 + q) * S + j) : 0ull;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t lane)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

// Random read of code k of lane j with a run-time bit width.
__device__ __forceinline__ uint64_t vbp_read(const uint64_t *__restrict__ base, uint32_t S, uint32_t j, uint64_t k,
                                             uint32_t bw)
{
    const uint64_t b = k * bw;
    const uint64_t w = b >> 6;
    const uint32_t off = (uint32_t)(b & 63);
    uint64_t v = gload(base + w * S + j) >> off;
    if (off + bw > 64) v |= gload(base + (w + 1) * S + j) << (64 - off);
    return bw == 64 ? v : (v & ((1ull << bw) - 1ull));
}

// ------------------------------------------------------------ wave ops
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, MS_WAVE);
    return x;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x |= __shfl_xor(x, o, MS_WAVE);
    return x;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Predicate of select as a value range [lo, hi], optionally negated:
// EQ [v,v]; NE !(EQ); LT [0,v-1]; LE [0,v]; GT [v+1,MAX]; GE [v,MAX].
struct ms_pred {
    uint64_t lo, hi;
    uint32_t neg;     // 1: select values outside [lo, hi]
    uint32_t empty;   // 1: [lo, hi] is empty
};

__host__ __device__ inline ms_pred make_pred(uint32_t cmp, uint64_t v)
{
    ms_pred p = {0, ~0ull, 0, 0};
    switch (cmp) {
    case MS_CMP_EQ: p.lo = v; p.hi = v; break;
    case MS_CMP_NE: p.lo = v; p.hi = v; p.neg = 1; break;
    case MS_CMP_LT: if (v == 0) p.empty = 1; else p.hi = v - 1; break;
    case MS_CMP_LE: p.hi = v; break;
    case MS_CMP_GT: if (v == ~0ull) p.empty = 1; else p.lo = v + 1; break;
    case MS_CMP_GE: p.lo = v; break;
    default: p.empty = 1; break;
    }
    return p;
}

__device__ __forceinline__ bool pred_eval(const ms_pred &p, uint64_t x)
{
    const bool in = !p.empty && (x - p.lo) <= (p.hi - p.lo);
    return in != (bool)p.neg;
}

// Block-level specialisation of a predicate for codes in [0, maxcode]
// offset by `ref` (x = ref + code, no wrap-around): the codes that pass are
// either none, all, [clo, clo+cspan] or everything outside that range.
enum { PM_NONE = 0, PM_ALL = 1, PM_IN = 2, PM_OUT = 3 };

struct ms_cpred {
    uint32_t mode;
    uint64_t clo, cspan;
};

__device__ __forceinline__ ms_cpred code_pred(const ms_pred &p, uint64_t ref, uint64_t maxcode)
{
    ms_cpred c = {PM_NONE, 0, 0};
    const uint64_t vmax = ref + maxcode;                 // caller guarantees no wrap
    uint64_t lo = p.lo, hi = p.hi;
    bool in_empty = p.empty || hi < ref || lo > vmax;
    if (!in_empty) {
        if (lo < ref) lo = ref;
        if (hi > vmax) hi = vmax;
    }
    const bool in_all = !in_empty && lo == ref && hi == vmax;
    if (!p.neg) {
        if (in_empty) c.mode = PM_NONE;
        else if (in_all) c.mode = PM_ALL;
        else { c.mode = PM_IN; c.clo = lo - ref; c.cspan = hi - lo; }
    } else {
        if (in_empty) c.mode = PM_ALL;
        else if (in_all) c.mode = PM_NONE;
        else { c.mode = PM_OUT; c.clo = lo - ref; c.cspan = hi - lo; }
    }
    return c;
}

}  // namespace ms
