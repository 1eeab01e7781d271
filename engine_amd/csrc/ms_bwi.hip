// ms_bwi.hip -- blockwise_individual_f (blockwise_individual.h:113-600).
//
// A column of this format is cut into blocks of B values.  Per block the
// image holds the block's format code (format_selector_commons.h:25-103),
// zero padding up to the next 64-byte boundary of the image (the
// create_aligned_ptr of the 8-lane processing style) and one morph_batch
// call of the code's default format over the B values: a self-contained
// image of B values in that format, called a view here.
//
// On the device a column of this format keeps, besides the image:
//   * per view its payload offset, size and code (host copy in ms_bwi);
//   * the block index entries of the dynamic_vbp / delta / for views,
//     grouped by code, as absolute image offsets, so that all views of one
//     code read as one virtual column of that code's format whose data
//     pointer is the image base (its chunk c is the c-th view of the code);
//   * for_f references and delta_f tile carries of those views, grouped
//     the same way (each view's carries restart at its first value);
//   * device meta data for every view and every group.
//
// Compression computes every view's payload size first (emit size passes
// over the views of a code, a walk per group_simple view, a run count per
// rle view, closed forms for uncompr_f and static_vbp_f), places the views
// with one scan, then packs every code's views with one launch each.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ms_units.hpp"

namespace ms {

__constant__ uint8_t c_bwi_num[15] = {64, 32, 21, 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1};
__constant__ uint8_t c_bwi_bits[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 21, 32, 64};

// The default formats of the 8-lane processing style (default_formats.h:
// 35-94).  Codes 68 and 70 name DEFAULT_DELTA_GROUP_SIMPLE_F and
// DEFAULT_FOR_GROUP_SIMPLE_F, which default_formats.h does not define.
bool bwi_inner(uint32_t code, ms_format_desc &d)
{
    d = ms_format_desc{};
    if (code == 0) { d.kind = MS_UNCOMPR; return true; }
    if (code <= 64) { d.kind = MS_STATIC_VBP; d.bw = code; d.step = 8; return true; }
    switch (code) {
    case MS_BWI_DYNAMIC_VBP: d.kind = MS_DYNAMIC_VBP; d.block = 512; d.page = 64; d.step = 8; return true;
    case MS_BWI_GROUP_SIMPLE: d.kind = MS_GROUP_SIMPLE; d.step = 8; return true;
    case MS_BWI_DELTA_DYNAMIC_VBP:
    case MS_BWI_FOR_DYNAMIC_VBP:
        d.kind = code == MS_BWI_DELTA_DYNAMIC_VBP ? MS_DELTA : MS_FOR;
        d.block = 1024;
        d.page = code == MS_BWI_DELTA_DYNAMIC_VBP ? 0 : 8;
        d.step = 8;
        d.inner_kind = MS_DYNAMIC_VBP;
        d.inner_block = 512;
        d.inner_page = 64;
        d.inner_step = 8;
        return true;
    case MS_BWI_RLE: d.kind = MS_RLE; return true;
    default: return false;
    }
}

// ---------------------------------------------------------------- placement
__global__ __launch_bounds__(WG) void k_bwi_padded(const uint64_t *psz, uint64_t nb, uint64_t *t)
{
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b <= nb; b += (uint64_t)gridDim.x * WG)
        t[b] = b < nb ? ms_align64(psz[b] + 1) : 0;
}

// q[b] = payload offset of view b (absolute; base0 = offset of the first
// code byte), from the exclusive scan of the padded sizes.
__global__ __launch_bounds__(WG) void k_bwi_offsets(uint64_t *q, uint64_t nb, uint64_t base0)
{
    const uint64_t first = ms_align64(base0 + 1);
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b <= nb; b += (uint64_t)gridDim.x * WG) q[b] += first;
}

__global__ __launch_bounds__(WG) void k_bwi_codes(uint8_t *out, const uint8_t *codes, const uint64_t *q,
                                                  const uint64_t *psz, uint64_t nb, uint64_t base0)
{
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * WG)
        out[b == 0 ? base0 : q[b - 1] + psz[b - 1]] = codes[b];
}

__global__ __launch_bounds__(WG) void k_bwi_cout(const uint64_t *q, const uint32_t *cmap, uint64_t n, uint64_t *cout)
{
    for (uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x; c < n; c += (uint64_t)gridDim.x * WG)
        cout[c] = q[cmap[c]];
}

__global__ __launch_bounds__(WG) void k_bwi_scatter_sizes(const uint64_t *csz, const uint32_t *cmap, uint64_t n,
                                                          uint64_t *psz)
{
    for (uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x; c < n; c += (uint64_t)gridDim.x * WG)
        psz[cmap[c]] = csz[c];
}

// ------------------------------------------------- uncompr_f and static_vbp_f
__global__ __launch_bounds__(WG) void k_bwi_plain_sizes(const uint8_t *codes, uint64_t nb, uint64_t B, uint64_t *psz)
{
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * WG) {
        const uint32_t c = codes[b];
        if (c <= 64) psz[b] = c == 0 ? 8 * B : c * B / 8;
    }
}

// Every output word of the uncompr_f / static_vbp_f<vbp_l<bw, 8>> views
// straight from its codes (vbp_routine_gen.py:101-111: codes ORed in
// unmasked).
__global__ __launch_bounds__(WG) void k_bwi_pack_plain(const uint64_t *v, uint64_t B, uint64_t nb,
                                                       const uint8_t *codes, const uint64_t *q, uint8_t *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nb * B; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = i / B, W = i - b * B;
        const uint32_t c = codes[b];
        if (c > 64) continue;
        const uint64_t *x = v + b * B;
        uint64_t *o = (uint64_t *)(out + q[b]);
        if (c == 0) { o[W] = x[W]; continue; }
        const uint64_t bw = c, words = B * bw / 64;
        if (W >= words) continue;
        const uint64_t j = W % 8, w = W / 8, b0 = 64 * w, b1 = b0 + 64;
        uint64_t k = (b0 + bw - 1) / bw;
        uint64_t word = 0;
        if (k > 0) {
            const uint64_t kc = k - 1;
            if (kc * bw + bw > b0) word |= x[kc * 8 + j] >> (b0 - kc * bw);
        }
        for (; k * bw < b1; k++) word |= x[k * 8 + j] << (k * bw - b0);
        o[W] = word;
    }
}

__global__ __launch_bounds__(WG) void k_bwi_unpack_plain(const uint8_t *img, uint64_t B, uint64_t nb,
                                                         const uint8_t *codes, const uint64_t *q, uint64_t *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nb * B; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = i / B, r = i - b * B;
        const uint32_t c = codes[b];
        if (c > 64) continue;
        const uint64_t *w = (const uint64_t *)(img + q[b]);
        out[i] = c == 0 ? w[r] : vbp_read(w, 8, (uint32_t)(r % 8), r / 8, c);
    }
}

__global__ __launch_bounds__(WG) void k_bwi_sum_plain(const uint8_t *img, uint64_t B, uint64_t nb,
                                                      const uint8_t *codes, const uint64_t *q, ms_pred p,
                                                      uint64_t *res)
{
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nb * B; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = i / B, r = i - b * B;
        const uint32_t c = codes[b];
        if (c > 64) continue;
        const uint64_t *w = (const uint64_t *)(img + q[b]);
        const uint64_t x = c == 0 ? w[r] : vbp_read(w, 8, (uint32_t)(r % 8), r / 8, c);
        if (pred_eval(p, x)) acc += x;
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd((unsigned long long *)res, (unsigned long long)acc);
}

// ------------------------------------------------------------ group_simple_f
// Bit width of every group (OR of its 8 values) of the listed blocks.
__global__ __launch_bounds__(WG) void k_bwi_gs_widths(const uint64_t *v, uint64_t B, const uint32_t *blocks,
                                                      uint64_t nblk, uint8_t *gbw)
{
    const uint64_t G = B / 8;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nblk * G; i += (uint64_t)gridDim.x * WG) {
        const uint64_t k = i / G, g = i - k * G;
        const uint64_t *x = v + (uint64_t)blocks[k] * B + 8 * g;
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) o |= x[j];
        gbw[i] = (uint8_t)(o ? 64 - __clzll(o) : 0);
    }
}

// The first mode (group_simple_routine_gen.py:101-168) whose next
// min(count, groups left, 64) groups all fit its width: lane l holds the
// width of group p + l.
__device__ __forceinline__ uint32_t gs_mode(uint32_t w, uint64_t avail)
{
    for (uint32_t i = 0; i < 14; i++) {
        const uint64_t lim = c_bwi_num[i] < avail ? c_bwi_num[i] : avail;
        const uint64_t lm = lim >= 64 ? ~0ull : ((1ull << lim) - 1ull);
        if ((__ballot(w > c_bwi_bits[i]) & lm) == 0) return i;
    }
    return 14;
}

// Size of one group_simple_f<8, uint64_t, 64> call over a block: a wave
// per listed block walks its selectors.
__global__ __launch_bounds__(64) void k_bwi_gs_walk(const uint8_t *gbw, uint64_t B, const uint32_t *blocks,
                                                    uint64_t nblk, uint64_t *psz, uint64_t *cs_out)
{
    const uint64_t G = B / 8;
    const uint32_t t = threadIdx.x;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint8_t *gb = gbw + k * G;
        uint64_t p = 0, cs = 0;
        while (p < G) {
            const uint64_t avail = G - p < 64 ? G - p : 64;
            const uint32_t w = t < avail ? gb[p + t] : 0;
            const uint32_t m = gs_mode(w, avail);
            if (c_bwi_num[m] > avail) break;
            cs++;
            p += c_bwi_num[m];
        }
        if (t == 0) {
            const uint64_t R = G - p;
            psz[blocks[k]] = ms_align64(9 + (cs + 1) / 2) + 64 * (cs + (R ? 1 : 0));
            cs_out[k] = cs;
        }
    }
}

// Writes the call: selector count, groups left, selectors (low nibble
// first), zero padding to 64 bytes, then 8 words per selector; lane
// 8 r + j ORs the groups r, r + 8, ... of lane j and the 8 partial words
// of lane j are ORed across the wave.
__global__ __launch_bounds__(64) void k_bwi_gs_pack(const uint64_t *v, const uint8_t *gbw, uint64_t B,
                                                    const uint32_t *blocks, uint64_t nblk, const uint64_t *cs_in,
                                                    const uint64_t *q, uint8_t *out)
{
    const uint64_t G = B / 8;
    const uint32_t t = threadIdx.x, j = t & 7, r = t >> 3;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint8_t *gb = gbw + k * G;
        const uint64_t *x = v + (uint64_t)blocks[k] * B;
        uint8_t *o = out + q[blocks[k]];
        const uint64_t cs = cs_in[k];
        uint64_t *data = (uint64_t *)(o + ms_align64(9 + (cs + 1) / 2));
        uint64_t p = 0, s = 0;
        uint32_t nib = 0;
        while (p < G) {
            const uint64_t avail = G - p < 64 ? G - p : 64;
            const uint32_t w = t < avail ? gb[p + t] : 0;
            const uint32_t m = gs_mode(w, avail);
            const uint32_t nn = c_bwi_num[m], bits = c_bwi_bits[m];
            if (nn > avail) break;
            uint64_t word = 0;
            for (uint32_t g = r; g < nn; g += 8) word |= x[(p + g) * 8 + j] << (g * bits);
            word |= __shfl_xor(word, 8, 64);
            word |= __shfl_xor(word, 16, 64);
            word |= __shfl_xor(word, 32, 64);
            if (t < 8) data[s * 8 + j] = word;
            if (t == 0) {
                if (s & 1) o[9 + s / 2] = (uint8_t)(nib | (m << 4));
                else nib = m;
            }
            s++;
            p += nn;
        }
        const uint64_t R = G - p;
        if (R) {
            const uint32_t bits = (uint32_t)(64 / R);
            uint64_t word = 0;
            for (uint32_t g = r; g < R; g += 8) word |= x[(p + g) * 8 + j] << (g * bits);
            word |= __shfl_xor(word, 8, 64);
            word |= __shfl_xor(word, 16, 64);
            word |= __shfl_xor(word, 32, 64);
            if (t < 8) data[s * 8 + j] = word;
        }
        if (t == 0) {
            if (s & 1) o[9 + s / 2] = (uint8_t)nib;
            *(uint64_t *)o = cs;
            o[8] = (uint8_t)R;
        }
    }
}

// Decodes listed group_simple views into out + (blocks[k] * B): a
// workgroup per view; the value offsets of its selectors by a running
// scan over chunks of WG selectors.
__global__ __launch_bounds__(WG) void k_bwi_gs_unpack(const uint8_t *img, uint64_t B, const uint32_t *blocks,
                                                      const uint64_t *q, uint64_t nblk, uint64_t *out, int compact)
{
    __shared__ uint64_t s_scan[WG];
    const uint32_t t = threadIdx.x;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint8_t *o = img + q[blocks[k]];
        uint64_t *dst = out + (compact ? k : (uint64_t)blocks[k]) * B;
        const uint64_t cs = *(const uint64_t *)o;
        const uint32_t R = o[8];
        const uint64_t *data = (const uint64_t *)(o + ms_align64(9 + (cs + 1) / 2));
        uint64_t carry = 0;
        for (uint64_t c0 = 0; c0 < cs; c0 += WG) {
            const uint64_t s = c0 + t;
            uint32_t m = 0, nn = 0;
            if (s < cs) {
                m = (o[9 + s / 2] >> (4 * (s & 1))) & 15;
                nn = m < 15 ? c_bwi_num[m] : 0;
            }
            s_scan[t] = nn;
            __syncthreads();
            for (uint32_t d = 1; d < WG; d <<= 1) {
                const uint64_t y = t >= d ? s_scan[t - d] : 0;
                __syncthreads();
                s_scan[t] += y;
                __syncthreads();
            }
            const uint64_t g0 = carry + s_scan[t] - nn;   // first group of selector s
            const uint64_t total = s_scan[WG - 1];
            __syncthreads();
            if (s < cs && m < 15 && (g0 + nn) * 8 <= B) {
                const uint32_t bits = c_bwi_bits[m];
                const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
                for (uint32_t j = 0; j < 8; j++) {
                    const uint64_t w = data[s * 8 + j];
                    for (uint32_t g = 0; g < nn; g++)   // the last group is not masked
                        dst[(g0 + g) * 8 + j] = g + 1 < nn ? (w >> (g * bits)) & mask : w >> (g * bits);
                }
            }
            carry += total;
        }
        if (R && R < 64 && (carry + R) * 8 <= B) {
            const uint32_t bits = 64 / R;
            const uint64_t mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
            for (uint32_t i = t; i < 8 * R; i += WG) {
                const uint32_t g = i / 8, j = i % 8;
                dst[(carry + g) * 8 + j] = (data[cs * 8 + j] >> (g * bits)) & mask;
            }
        }
    }
}

// --------------------------------------------------------------------- rle_f
__global__ __launch_bounds__(64) void k_bwi_rle_runs(const uint64_t *v, uint64_t B, const uint32_t *blocks,
                                                     uint64_t nblk, uint64_t *psz)
{
    const uint32_t t = threadIdx.x;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint64_t *x = v + (uint64_t)blocks[k] * B;
        uint64_t runs = 0;
        for (uint64_t i0 = 0; i0 < B; i0 += 64) {
            const uint64_t i = i0 + t;
            const bool st = i < B && (i == 0 || x[i] != x[i - 1]);
            runs += __popcll(__ballot(st));
        }
        if (t == 0) psz[blocks[k]] = 16 * runs;
    }
}

// (value, length) pairs of the maximal runs of a block (rle.h:95-147):
// a start writes its value and the length of the run before it.
__global__ __launch_bounds__(64) void k_bwi_rle_pack(const uint64_t *v, uint64_t B, const uint32_t *blocks,
                                                     uint64_t nblk, const uint64_t *q, uint8_t *out)
{
    const uint32_t t = threadIdx.x;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint64_t *x = v + (uint64_t)blocks[k] * B;
        uint64_t *o = (uint64_t *)(out + q[blocks[k]]);
        uint64_t runs = 0, last = 0;   // runs so far, start of the latest run
        for (uint64_t i0 = 0; i0 < B; i0 += 64) {
            const uint64_t i = i0 + t;
            const bool st = i < B && (i == 0 || x[i] != x[i - 1]);
            const uint64_t m = __ballot(st);
            if (st) {
                const uint64_t below = m & ((1ull << t) - 1ull);
                const uint64_t r = runs + __popcll(below);
                const uint64_t prev = below ? i0 + 63 - __clzll(below) : last;
                o[2 * r] = x[i];
                if (r) o[2 * (r - 1) + 1] = i - prev;
            }
            if (m) last = i0 + 63 - __clzll(m);
            runs += __popcll(m);
        }
        if (t == 0 && runs) o[2 * (runs - 1) + 1] = B - last;
    }
}

// Expands listed rle views into out + blocks[k] * B: a workgroup per view,
// the run starts by a running scan over chunks of WG runs.
__global__ __launch_bounds__(WG) void k_bwi_rle_expand(const uint8_t *img, uint64_t B, const uint32_t *blocks,
                                                       const uint64_t *q, const uint64_t *psz, uint64_t nblk,
                                                       uint64_t *out)
{
    __shared__ uint64_t s_scan[WG];
    const uint32_t t = threadIdx.x;
    for (uint64_t k = blockIdx.x; k < nblk; k += gridDim.x) {
        const uint64_t *o = (const uint64_t *)(img + q[blocks[k]]);
        const uint64_t runs = psz[blocks[k]] / 16;
        uint64_t *dst = out + (uint64_t)blocks[k] * B;
        uint64_t carry = 0;
        for (uint64_t r0 = 0; r0 < runs; r0 += WG) {
            const uint64_t r = r0 + t;
            const uint64_t len = r < runs ? o[2 * r + 1] : 0;
            s_scan[t] = len;
            __syncthreads();
            for (uint32_t d = 1; d < WG; d <<= 1) {
                const uint64_t y = t >= d ? s_scan[t - d] : 0;
                __syncthreads();
                s_scan[t] += y;
                __syncthreads();
            }
            const uint64_t start = carry + s_scan[t] - len;
            const uint64_t total = s_scan[WG - 1];
            __syncthreads();
            if (r < runs && start <= B && len <= B - start) {
                const uint64_t val = o[2 * r];
                for (uint64_t i = 0; i < len; i++) dst[start + i] = val;
            }
            carry += total;
        }
    }
}

// Sum of the selected values of listed rle views: (view, run) pairs
// grid-strided.
__global__ __launch_bounds__(WG) void k_bwi_rle_sum(const uint8_t *img, const uint32_t *blocks, const uint64_t *q,
                                                    const uint64_t *psz, uint64_t nblk, ms_pred p, uint64_t *res)
{
    uint64_t acc = 0;
    for (uint64_t k = blockIdx.y; k < nblk; k += gridDim.y) {
        const uint64_t *o = (const uint64_t *)(img + q[blocks[k]]);
        const uint64_t runs = psz[blocks[k]] / 16;
        for (uint64_t r = (uint64_t)blockIdx.x * WG + threadIdx.x; r < runs; r += (uint64_t)gridDim.x * WG) {
            const uint64_t x = o[2 * r];
            if (pred_eval(p, x)) acc += x * o[2 * r + 1];
        }
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd((unsigned long long *)res, (unsigned long long)acc);
}

// ------------------------------------------------------------------- delta
// Tile carries of the delta views: per view, per 2048-value tile, the 8
// values before the tile (zero for its first tile: every view's delta
// starts from zero).
__global__ __launch_bounds__(WG) void k_bwi_carries(const uint64_t *v, uint64_t B, const uint32_t *cmap, uint64_t n,
                                                    uint64_t T, uint64_t *carries)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n * T * 8; i += (uint64_t)gridDim.x * WG) {
        const uint64_t c = i / (T * 8), rem = i - c * T * 8, tile = rem / 8, j = rem % 8;
        carries[i] = tile == 0 ? 0ull : v[(uint64_t)cmap[c] * B + tile * MS_DELTA_TILE - 8 + j];
    }
}

// out[cmap[c] * B + i] = tmp[c * B + i]
__global__ __launch_bounds__(WG) void k_bwi_scatter(const uint64_t *tmp, const uint32_t *cmap, uint64_t n, uint64_t B,
                                                    uint64_t *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n * B; i += (uint64_t)gridDim.x * WG) {
        const uint64_t c = i / B;
        out[(uint64_t)cmap[c] * B + (i - c * B)] = tmp[i];
    }
}

__global__ __launch_bounds__(WG) void k_bwi_sum_values(const uint64_t *x, uint64_t n, ms_pred p, uint64_t *res)
{
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        if (pred_eval(p, x[i])) acc += x[i];
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd((unsigned long long *)res, (unsigned long long)acc);
}

// ---------------------------------------------------------------- selector
// exhaustive_size_format_selector.h:80-141: static_vbp_f at the block's
// maximum bit width in closed form (best, below 64 bits) or uncompr_f,
// then dynamic_vbp, group_simple, delta + dynamic_vbp and for +
// dynamic_vbp in this order, each replacing the choice when it is smaller
// (best) / not smaller (worst).
__global__ __launch_bounds__(WG) void k_bwi_choose(const uint32_t *bws, uint64_t spb, const uint64_t *s_dvbp,
                                                   const uint64_t *s_gs, const uint64_t *s_delta,
                                                   const uint64_t *s_for, uint64_t nb, uint64_t B, uint32_t best,
                                                   uint8_t *codes)
{
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * WG) {
        uint32_t mbw = 1;
        for (uint64_t s = 0; s < spb; s++) mbw = bws[b * spb + s] > mbw ? bws[b * spb + s] : mbw;
        uint64_t size;
        uint32_t code;
        if (best && mbw < 64) { size = (uint64_t)mbw * B / 8; code = mbw; }
        else { size = 8 * B; code = 0; }
        const uint64_t cand[4] = {s_dvbp[b], s_gs[b], s_delta[b], s_for[b]};
        const uint32_t cc[4] = {MS_BWI_DYNAMIC_VBP, MS_BWI_GROUP_SIMPLE, MS_BWI_DELTA_DYNAMIC_VBP,
                                MS_BWI_FOR_DYNAMIC_VBP};
        for (int i = 0; i < 4; i++)
            if ((best != 0) == (cand[i] < size)) { size = cand[i]; code = cc[i]; }
        codes[b] = (uint8_t)code;
    }
}

// ======================================================================
// Host side
// ======================================================================
static uint64_t bwi_T(uint64_t B) { return (B + MS_DELTA_TILE - 1) / MS_DELTA_TILE; }

static int kind_of(uint32_t code)
{
    switch (code) {
    case MS_BWI_DYNAMIC_VBP: return BWK_DVBP;
    case MS_BWI_DELTA_DYNAMIC_VBP: return BWK_DELTA;
    case MS_BWI_FOR_DYNAMIC_VBP: return BWK_FOR;
    case MS_BWI_GROUP_SIMPLE: return BWK_GS;
    case MS_BWI_RLE: return BWK_RLE;
    default: return BWK_PLAIN;
    }
}

static const uint32_t k_group_code[3] = {MS_BWI_DYNAMIC_VBP, MS_BWI_DELTA_DYNAMIC_VBP, MS_BWI_FOR_DYNAMIC_VBP};

uint64_t bwi_tab_bytes(uint64_t nb)
{
    return (nb + 3) * sizeof(ms_dev_meta) + nb * 16 + (nb * 4 + 7) / 8 * 8 + (nb + 7) / 8 * 8;
}

bwi_dev_tab bwi_tab(const ms_col *c)
{
    bwi_dev_tab t;
    const uint64_t nb = c->bwi ? c->bwi->nb : 0;
    uint8_t *p = (uint8_t *)c->vmeta;
    t.meta = (ms_dev_meta *)p;
    p += (nb + 3) * sizeof(ms_dev_meta);
    t.q = (uint64_t *)p;
    p += nb * 8;
    t.psz = (uint64_t *)p;
    p += nb * 8;
    t.lists = (uint32_t *)p;
    p += (nb * 4 + 7) / 8 * 8;
    t.codes = p;
    return t;
}

// Index entries and reference slots of the views of codes 65, 67, 69 and
// the lists, from the codes.
static void bwi_layout(ms_bwi &h, const uint8_t *codes, std::vector<uint32_t> &lists)
{
    const uint64_t nb = h.nb, B = h.B;
    for (int k = 0; k < BWK_N; k++) h.lcount[k] = 0;
    for (uint64_t b = 0; b < nb; b++) h.lcount[kind_of(codes[b])]++;
    uint64_t o = 0;
    for (int k = 0; k < BWK_N; k++) { h.loff[k] = o; o += h.lcount[k]; }
    lists.assign(nb, 0);
    uint64_t fill[BWK_N] = {0, 0, 0, 0, 0, 0};
    h.v.assign(nb, ms_bwi_view{});
    for (uint64_t b = 0; b < nb; b++) {
        const int k = kind_of(codes[b]);
        h.v[b].code = codes[b];
        h.v[b].slot = (uint32_t)fill[k];
        lists[h.loff[k] + fill[k]++] = (uint32_t)b;
    }
    h.idx_off[BWK_DVBP] = 0;
    h.idx_off[BWK_DELTA] = h.lcount[BWK_DVBP] * (B / 512);
    h.idx_off[BWK_FOR] = h.idx_off[BWK_DELTA] + h.lcount[BWK_DELTA] * (B / 512);
    h.refs_off[BWK_DVBP] = 0;
    h.refs_off[BWK_DELTA] = 0;
    // 8 slots of slack behind the delta carries: a decoding pass over a
    // view (launch_delta_carries) writes one tile more than the view has
    // when B is a multiple of the tile, into the next view's first tile,
    // which that view's own pass then rewrites
    h.refs_off[BWK_FOR] = h.lcount[BWK_DELTA] ? h.lcount[BWK_DELTA] * bwi_T(B) * 8 + 8 : 0;
}

uint64_t bwi_index_entries(const ms_bwi &h) { return h.idx_off[BWK_FOR] + h.lcount[BWK_FOR] * (h.B / 512); }
uint64_t bwi_refs_entries(const ms_bwi &h) { return h.refs_off[BWK_FOR] + h.lcount[BWK_FOR] * (h.B / 1024); }

static ms_format_desc inner_of_code(uint32_t code)
{
    ms_format_desc d;
    bwi_inner(code, d);
    return d;
}

bool bwi_code_ok(uint32_t code)
{
    ms_format_desc d;
    return bwi_inner(code, d);
}

int bwi_plan(const ms_format_desc &f, const uint64_t *d_v, uint64_t n, const uint8_t *h_codes, uint64_t base0,
             bwi_work &w, temps &sc)
{
    hipStream_t s = sc.s;
    const uint64_t B = f.block, nb = n / B;
    w.v = d_v;
    w.h = ms_bwi{};
    w.h.n = n;
    w.h.nb = nb;
    w.h.B = B;
    w.h.C = 0;
    w.base0 = base0;
    w.codes.assign(nb, 0);
    if (nb == 0) {
        bwi_layout(w.h, w.codes.data(), w.lists);
        return MS_OK;
    }
    if ((uint64_t)(uint32_t)nb != nb) return ms_fail(MS_ERR_CAPACITY, "more than 2^32 blocks");
    w.d_codes = (uint8_t *)sc.get(nb + 64);
    w.d_lists = (uint32_t *)sc.get(nb * 4 + 64);
    w.d_psz = (uint64_t *)sc.get((nb + 1) * 8 + 64);
    w.d_q = (uint64_t *)sc.get((nb + 1) * 8 + 64);
    uint64_t *t = (uint64_t *)sc.get((nb + 1) * 8 + 64);
    if (!w.d_codes || !w.d_lists || !w.d_psz || !w.d_q || !t) return ms_fail(MS_ERR_NOMEM, "bwi scratch");
    const uint64_t G = B / 8;
    if (h_codes || f.bw == MS_BWI_SEL_STATIC) {
        for (uint64_t b = 0; b < nb; b++) {
            const uint32_t c = h_codes ? h_codes[b] : f.page;
            if (!bwi_code_ok(c)) return ms_fail(MS_ERR_INVALID, "format code %u of block %llu", c, (unsigned long long)b);
            w.codes[b] = (uint8_t)c;
        }
        MS_HIP(hipMemcpyAsync(w.d_codes, w.codes.data(), nb, hipMemcpyHostToDevice, s));
    } else {
        // the exhaustive selector: every candidate's size of every block
        std::vector<uint32_t> ident(nb);
        for (uint64_t b = 0; b < nb; b++) ident[b] = (uint32_t)b;
        uint32_t *d_ident = (uint32_t *)sc.get(nb * 4 + 64);
        uint64_t *sz[3], *s_gs = (uint64_t *)sc.get(nb * 8 + 64), *cs = (uint64_t *)sc.get(nb * 8 + 64);
        uint8_t *gbw = (uint8_t *)sc.get(nb * G + 64);
        if (!d_ident || !s_gs || !cs || !gbw) return ms_fail(MS_ERR_NOMEM, "bwi selector scratch");
        MS_HIP(hipMemcpyAsync(d_ident, ident.data(), nb * 4, hipMemcpyHostToDevice, s));
        chunk_emit ce[3];
        for (int k = 0; k < 3; k++) {
            sz[k] = (uint64_t *)sc.get(nb * 8 + 64);
            if (!sz[k]) return ms_fail(MS_ERR_NOMEM, "bwi selector scratch");
            ce[k] = chunk_emit{};
            ce[k].f = inner_of_code(k_group_code[k]);
            ce[k].v = d_v;
            ce[k].cmap = d_ident;
            ce[k].nchunks = nb;
            ce[k].K = B;
            MS_TRY(launch_chunk_sizes(ce[k], sz[k], sc));
        }
        hipLaunchKernelGGL(k_bwi_gs_widths, dim3(grid_for(nb * G, WG * 4ull)), dim3(WG), 0, s, d_v, B,
                           (const uint32_t *)d_ident, nb, gbw);
        hipLaunchKernelGGL(k_bwi_gs_walk, dim3(grid_for(nb, 1, 65536)), dim3(64), 0, s, (const uint8_t *)gbw, B,
                           (const uint32_t *)d_ident, nb, s_gs, cs);
        hipLaunchKernelGGL(k_bwi_choose, dim3(grid_for(nb, WG)), dim3(WG), 0, s, (const uint32_t *)ce[0].bws,
                           B / 512, (const uint64_t *)sz[0], (const uint64_t *)s_gs, (const uint64_t *)sz[1],
                           (const uint64_t *)sz[2], nb, B, (uint32_t)(f.bw == MS_BWI_SEL_BEST), w.d_codes);
        MS_HIP(hipGetLastError());
        MS_HIP(hipMemcpyAsync(w.codes.data(), w.d_codes, nb, hipMemcpyDeviceToHost, s));
        MS_HIP(hipStreamSynchronize(s));
    }
    bwi_layout(w.h, w.codes.data(), w.lists);
    MS_HIP(hipMemcpyAsync(w.d_lists, w.lists.data(), nb * 4, hipMemcpyHostToDevice, s));
    const ms_bwi &h = w.h;
    // payload sizes
    hipLaunchKernelGGL(k_bwi_plain_sizes, dim3(grid_for(nb, WG)), dim3(WG), 0, s, (const uint8_t *)w.d_codes, nb, B,
                       w.d_psz);
    for (int k = 0; k < 3; k++) {
        w.ce[k] = chunk_emit{};
        if (!h.lcount[k]) continue;
        const uint32_t *list = w.d_lists + h.loff[k];
        uint64_t *csz = (uint64_t *)sc.get(h.lcount[k] * 8 + 64);
        if (!csz) return ms_fail(MS_ERR_NOMEM, "bwi scratch");
        w.ce[k].f = inner_of_code(k_group_code[k]);
        w.ce[k].v = d_v;
        w.ce[k].cmap = list;
        w.ce[k].nchunks = h.lcount[k];
        w.ce[k].K = B;
        MS_TRY(launch_chunk_sizes(w.ce[k], csz, sc));
        hipLaunchKernelGGL(k_bwi_scatter_sizes, dim3(grid_for(h.lcount[k], WG)), dim3(WG), 0, s,
                           (const uint64_t *)csz, list, h.lcount[k], w.d_psz);
    }
    w.gbw = nullptr;
    w.gs_cs = nullptr;
    if (h.lcount[BWK_GS]) {
        const uint64_t m = h.lcount[BWK_GS];
        const uint32_t *list = w.d_lists + h.loff[BWK_GS];
        w.gbw = (uint8_t *)sc.get(m * G + 64);
        w.gs_cs = (uint64_t *)sc.get(m * 8 + 64);
        if (!w.gbw || !w.gs_cs) return ms_fail(MS_ERR_NOMEM, "bwi scratch");
        hipLaunchKernelGGL(k_bwi_gs_widths, dim3(grid_for(m * G, WG * 4ull)), dim3(WG), 0, s, d_v, B, list, m, w.gbw);
        hipLaunchKernelGGL(k_bwi_gs_walk, dim3(grid_for(m, 1, 65536)), dim3(64), 0, s, (const uint8_t *)w.gbw, B, list,
                           m, w.d_psz, w.gs_cs);
    }
    if (h.lcount[BWK_RLE]) {
        const uint64_t m = h.lcount[BWK_RLE];
        hipLaunchKernelGGL(k_bwi_rle_runs, dim3(grid_for(m, 1, 65536)), dim3(64), 0, s, d_v, B,
                           (const uint32_t *)(w.d_lists + h.loff[BWK_RLE]), m, w.d_psz);
    }
    // placement
    hipLaunchKernelGGL(k_bwi_padded, dim3(grid_for(nb + 1, WG)), dim3(WG), 0, s, (const uint64_t *)w.d_psz, nb, t);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan(t, w.d_q, nb + 1, sc));
    hipLaunchKernelGGL(k_bwi_offsets, dim3(grid_for(nb + 1, WG)), dim3(WG), 0, s, w.d_q, nb, base0);
    MS_HIP(hipGetLastError());
    std::vector<uint64_t> q(nb), psz(nb);
    MS_HIP(hipMemcpyAsync(q.data(), w.d_q, nb * 8, hipMemcpyDeviceToHost, s));
    MS_HIP(hipMemcpyAsync(psz.data(), w.d_psz, nb * 8, hipMemcpyDeviceToHost, s));
    MS_HIP(hipStreamSynchronize(s));
    for (uint64_t b = 0; b < nb; b++) {
        w.h.v[b].off = q[b];
        w.h.v[b].size = psz[b];
    }
    w.h.C = q[nb - 1] + psz[nb - 1];
    return MS_OK;
}

// Host copy of the view table into the column's device table.
static int upload_tab(ms_col *c, const ms_bwi &h, const std::vector<uint8_t> &codes,
                      const std::vector<uint32_t> &lists, hipStream_t s)
{
    const uint64_t nb = h.nb;
    const bwi_dev_tab t = bwi_tab(c);
    std::vector<ms_dev_meta> vm(nb + 3);
    std::vector<uint64_t> qs(2 * nb);
    for (uint64_t b = 0; b < nb; b++) {
        vm[b] = ms_dev_meta{};
        vm[b].count = h.B;
        qs[b] = h.v[b].off;
        qs[nb + b] = h.v[b].size;
    }
    for (int k = 0; k < 3; k++) {
        vm[nb + k] = ms_dev_meta{};
        vm[nb + k].count = h.lcount[k] * h.B;
    }
    MS_HIP(hipMemcpyAsync(t.meta, vm.data(), (nb + 3) * sizeof(ms_dev_meta), hipMemcpyHostToDevice, s));
    if (nb) {
        MS_HIP(hipMemcpyAsync(t.q, qs.data(), 16 * nb, hipMemcpyHostToDevice, s));
        MS_HIP(hipMemcpyAsync(t.lists, lists.data(), 4 * nb, hipMemcpyHostToDevice, s));
        MS_HIP(hipMemcpyAsync(t.codes, codes.data(), nb, hipMemcpyHostToDevice, s));
    }
    MS_HIP(hipStreamSynchronize(s));   // the host vectors die here
    return MS_OK;
}

__global__ __launch_bounds__(WG) void k_fill_u8(uint8_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) p[i] = 0;
}

int bwi_write_image(bwi_work &w, uint8_t *img, uint64_t cap, uint64_t *status, uint64_t *index_out,
                    uint64_t *refs_out, temps &sc)
{
    hipStream_t s = sc.s;
    const ms_bwi &h = w.h;
    const uint64_t nb = h.nb, B = h.B;
    if (!nb) return MS_OK;
    if (h.C > cap) return ms_fail(MS_ERR_CAPACITY, "blockwise_individual_f image of %llu bytes exceeds %llu",
                                  (unsigned long long)h.C, (unsigned long long)cap);
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(h.C - w.base0, WG * 16ull)), dim3(WG), 0, s, img + w.base0,
                       h.C - w.base0);
    hipLaunchKernelGGL(k_bwi_codes, dim3(grid_for(nb, WG)), dim3(WG), 0, s, img, (const uint8_t *)w.d_codes,
                       (const uint64_t *)w.d_q, (const uint64_t *)w.d_psz, nb, w.base0);
    if (h.lcount[BWK_PLAIN])
        hipLaunchKernelGGL(k_bwi_pack_plain, dim3(grid_for(nb * B, WG * 8ull)), dim3(WG), 0, s, w.v, B, nb,
                           (const uint8_t *)w.d_codes, (const uint64_t *)w.d_q, img);
    MS_HIP(hipGetLastError());
    for (int k = 0; k < 3; k++) {
        if (!h.lcount[k]) continue;
        const uint32_t *list = w.d_lists + h.loff[k];
        uint64_t *cout = (uint64_t *)sc.get(h.lcount[k] * 8 + 64);
        if (!cout) return ms_fail(MS_ERR_NOMEM, "bwi scratch");
        hipLaunchKernelGGL(k_bwi_cout, dim3(grid_for(h.lcount[k], WG)), dim3(WG), 0, s, (const uint64_t *)w.d_q, list,
                           h.lcount[k], cout);
        MS_TRY(launch_chunk_pack(w.ce[k], cout, img, cap, status, index_out + h.idx_off[k],
                                 k == BWK_FOR ? refs_out + h.refs_off[k] : nullptr, s));
        if (k == BWK_DELTA)
            hipLaunchKernelGGL(k_bwi_carries, dim3(grid_for(h.lcount[k] * bwi_T(B) * 8, WG)), dim3(WG), 0, s, w.v,
                               B, list, h.lcount[k], bwi_T(B), refs_out + h.refs_off[k]);
    }
    if (h.lcount[BWK_GS])
        hipLaunchKernelGGL(k_bwi_gs_pack, dim3(grid_for(h.lcount[BWK_GS], 1, 65536)), dim3(64), 0, s, w.v,
                           (const uint8_t *)w.gbw, B, (const uint32_t *)(w.d_lists + h.loff[BWK_GS]),
                           h.lcount[BWK_GS], (const uint64_t *)w.gs_cs, (const uint64_t *)w.d_q, img);
    if (h.lcount[BWK_RLE])
        hipLaunchKernelGGL(k_bwi_rle_pack, dim3(grid_for(h.lcount[BWK_RLE], 1, 65536)), dim3(64), 0, s, w.v, B,
                           (const uint32_t *)(w.d_lists + h.loff[BWK_RLE]), h.lcount[BWK_RLE],
                           (const uint64_t *)w.d_q, img);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

int bwi_pack(bwi_work &w, ms_col *c, temps &sc)
{
    hipStream_t s = sc.s;
    const ms_bwi &h = w.h;
    const uint64_t nb = h.nb, B = h.B;
    MS_TRY(bwi_write_image(w, c->data, c->size_alloc, &c->meta->status, c->index, c->refs, sc));
    const uint64_t rest = h.n - nb * B;
    const uint64_t R = ms_align64(h.C + c->g.gap);
    if (rest) MS_HIP(hipMemcpyAsync(c->data + R, w.v + nb * B, 8 * rest, hipMemcpyDeviceToDevice, s));
    ms_dev_meta dm = {};
    dm.count = h.n;
    dm.size_compr = h.C;
    dm.size_used = rest ? R + 8 * rest : h.C;
    MS_HIP(hipMemcpyAsync(c->meta, &dm, sizeof dm, hipMemcpyHostToDevice, s));
    c->bwi = new (std::nothrow) ms_bwi(h);
    if (!c->bwi) return ms_fail(MS_ERR_NOMEM, "host allocation");
    return upload_tab(c, h, w.codes, w.lists, s);   // synchronises (dm too)
}

// A temporary column over the views of one of codes 65 / 67 / 69 (all of
// them, or one: slot `one`).
static ms_col group_col(const ms_col *c, int k, int64_t one = -1)
{
    const ms_bwi &h = *c->bwi;
    const bwi_dev_tab t = bwi_tab(c);
    ms_col g = ms_col();
    g.f = inner_of_code(k_group_code[k]);
    g.g = ms_make_geom(g.f);
    g.data = c->data;
    g.stream = c->stream;
    const uint64_t B = h.B;
    if (one < 0) {
        g.meta = t.meta + h.nb + k;
        g.count_bound = h.lcount[k] * B;
        g.index = c->index + h.idx_off[k];
        g.refs = k == BWK_DVBP ? nullptr : c->refs + h.refs_off[k];
    } else {
        g.meta = t.meta;   // any view's meta: count B, no rest
        g.count_bound = B;
        g.index = c->index + h.idx_off[k] + (uint64_t)one * (B / 512);
        g.refs = k == BWK_DELTA ? c->refs + h.refs_off[k] + (uint64_t)one * bwi_T(B) * 8
               : k == BWK_FOR   ? c->refs + h.refs_off[k] + (uint64_t)one * (B / 1024)
                                : nullptr;
    }
    g.size_alloc = c->size_alloc;
    g.index_cap = g.count_bound / 512 + 1;
    g.refs_cap = k == BWK_DVBP ? 0 : k == BWK_FOR ? g.count_bound / 1024 + 1 : (g.count_bound / MS_DELTA_TILE + 1) * 8;
    return g;
}

int launch_bwi_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s)
{
    if (!in->bwi) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f column without views");
    const ms_bwi &h = *in->bwi;
    const bwi_dev_tab t = bwi_tab(in);
    const uint64_t nb = h.nb, B = h.B;
    temps sc(s);
    if (h.lcount[BWK_PLAIN])
        hipLaunchKernelGGL(k_bwi_unpack_plain, dim3(grid_for(nb * B, WG * 8ull)), dim3(WG), 0, s,
                           (const uint8_t *)in->data, B, nb, (const uint8_t *)t.codes, (const uint64_t *)t.q, d_out);
    MS_HIP(hipGetLastError());
    for (int k = 0; k < 3; k++) {
        const uint64_t m = h.lcount[k];
        if (!m) continue;
        const uint32_t *list = t.lists + h.loff[k];
        if (k == BWK_DELTA && B % MS_DELTA_TILE) {
            // tiles would straddle views: one view at a time
            for (uint64_t b = 0; b < nb; b++) {
                if (h.v[b].code != MS_BWI_DELTA_DYNAMIC_VBP) continue;
                const ms_col g = group_col(in, k, (int64_t)h.v[b].slot);
                MS_TRY(launch_decompress(&g, d_out + b * B, s));
            }
            continue;
        }
        uint64_t *tmp = (uint64_t *)sc.get(m * B * 8 + 64);
        if (!tmp) return ms_fail(MS_ERR_NOMEM, "bwi decompression scratch");
        const ms_col g = group_col(in, k);
        MS_TRY(launch_decompress(&g, tmp, s));
        hipLaunchKernelGGL(k_bwi_scatter, dim3(grid_for(m * B, WG * 8ull)), dim3(WG), 0, s, (const uint64_t *)tmp,
                           list, m, B, d_out);
        MS_HIP(hipGetLastError());
    }
    if (h.lcount[BWK_GS])
        hipLaunchKernelGGL(k_bwi_gs_unpack, dim3(grid_for(h.lcount[BWK_GS], 1, 65536)), dim3(WG), 0, s,
                           (const uint8_t *)in->data, B, (const uint32_t *)(t.lists + h.loff[BWK_GS]),
                           (const uint64_t *)t.q, h.lcount[BWK_GS], d_out, 0);
    if (h.lcount[BWK_RLE])
        hipLaunchKernelGGL(k_bwi_rle_expand, dim3(grid_for(h.lcount[BWK_RLE], 1, 65536)), dim3(WG), 0, s,
                           (const uint8_t *)in->data, B, (const uint32_t *)(t.lists + h.loff[BWK_RLE]),
                           (const uint64_t *)t.q, (const uint64_t *)t.psz, h.lcount[BWK_RLE], d_out);
    MS_HIP(hipGetLastError());
    const uint64_t rest = h.n - nb * B;
    if (rest)
        MS_HIP(hipMemcpyAsync(d_out + nb * B, in->data + ms_align64(h.C + in->g.gap), 8 * rest,
                              hipMemcpyDeviceToDevice, s));
    return MS_OK;
}

__global__ void k_bwi_fold(const uint64_t *slots, uint32_t n, uint64_t *res)
{
    uint64_t a = 0;
    for (uint32_t i = 0; i < n; i++) a += slots[i];
    *res = a;
}

int launch_bwi_select_sum(const ms_col *in, const ms_pred &p, uint64_t *d_result, bool accumulate, hipStream_t s)
{
    (void)accumulate;
    if (!in->bwi) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f column without views");
    const ms_bwi &h = *in->bwi;
    const bwi_dev_tab t = bwi_tab(in);
    const uint64_t nb = h.nb, B = h.B;
    const bool per_view_delta = B % MS_DELTA_TILE != 0;
    const uint32_t nslots = 8 + (per_view_delta ? (uint32_t)h.lcount[BWK_DELTA] : 0);
    temps sc(s);
    uint64_t *slots = (uint64_t *)sc.get(nslots * 8ull + 64);
    if (!slots) return ms_fail(MS_ERR_NOMEM, "bwi select_sum scratch");
    MS_TRY(launch_fill_u64(slots, nslots, 0, s));
    uint32_t si = 0;
    if (h.lcount[BWK_PLAIN])
        hipLaunchKernelGGL(k_bwi_sum_plain, dim3(grid_for(nb * B, WG * 16ull, 2048)), dim3(WG), 0, s,
                           (const uint8_t *)in->data, B, nb, (const uint8_t *)t.codes, (const uint64_t *)t.q, p,
                           slots + si);
    si++;
    MS_HIP(hipGetLastError());
    for (int k = 0; k < 3; k++) {
        const uint64_t m = h.lcount[k];
        if (!m) { si++; continue; }
        if (k == BWK_DELTA && per_view_delta) {
            for (uint64_t j = 0; j < m; j++) {
                const ms_col g = group_col(in, k, (int64_t)j);
                MS_TRY(launch_select_sum(&g, p, slots + 8 + j, false, s));
            }
            si++;
            continue;
        }
        const ms_col g = group_col(in, k);
        MS_TRY(launch_select_sum(&g, p, slots + si, false, s));
        si++;
    }
    if (h.lcount[BWK_GS]) {
        const uint64_t m = h.lcount[BWK_GS];
        uint64_t *tmp = (uint64_t *)sc.get(m * B * 8 + 64);
        if (!tmp) return ms_fail(MS_ERR_NOMEM, "bwi select_sum scratch");
        hipLaunchKernelGGL(k_bwi_gs_unpack, dim3(grid_for(m, 1, 65536)), dim3(WG), 0, s, (const uint8_t *)in->data, B,
                           (const uint32_t *)(t.lists + h.loff[BWK_GS]), (const uint64_t *)t.q, m, tmp, 1);
        hipLaunchKernelGGL(k_bwi_sum_values, dim3(grid_for(m * B, WG * 16ull, 2048)), dim3(WG), 0, s,
                           (const uint64_t *)tmp, m * B, p, slots + si);
    }
    si++;
    if (h.lcount[BWK_RLE])
        hipLaunchKernelGGL(k_bwi_rle_sum, dim3(64, (uint32_t)std::min<uint64_t>(h.lcount[BWK_RLE], 65535)), dim3(WG), 0,
                           s, (const uint8_t *)in->data, (const uint32_t *)(t.lists + h.loff[BWK_RLE]),
                           (const uint64_t *)t.q, (const uint64_t *)t.psz, h.lcount[BWK_RLE], p, slots + si);
    si++;
    const uint64_t rest = h.n - nb * B;
    if (rest)
        hipLaunchKernelGGL(k_bwi_sum_values, dim3(grid_for(rest, WG * 16ull, 2048)), dim3(WG), 0, s,
                           (const uint64_t *)(in->data + ms_align64(h.C + in->g.gap)), rest, p, slots + si);
    si++;
    MS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bwi_fold, dim3(1), dim3(1), 0, s, (const uint64_t *)slots, nslots, d_result);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ---------------------------------------------------------------- upload
// Walks a host image: per block the code byte, the padding and the view
// (its size from the view's own walk); index entries (absolute) and for_f
// references of codes 65 / 67 / 69 in kind order.
int bwi_walk_image(const ms_format_desc &f, const uint8_t *img, uint64_t count, uint64_t size_compr, bwi_host &out,
                   int (*walk)(const ms_format_desc &, const uint8_t *, uint64_t, uint64_t, std::vector<uint64_t> &,
                               std::vector<uint64_t> &, uint64_t *))
{
    const uint64_t B = f.block, nb = count / B;
    out.h = ms_bwi{};
    out.h.n = count;
    out.h.nb = nb;
    out.h.B = B;
    out.codes.assign(nb, 0);
    std::vector<uint64_t> off(nb), size(nb);
    std::vector<std::vector<uint64_t>> idx(3), refs(3);
    uint64_t pos = 0;
    for (uint64_t b = 0; b < nb; b++) {
        if (pos >= size_compr) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: block %llu past sizeCompr",
                                              (unsigned long long)b);
        const uint32_t code = img[pos];
        ms_format_desc d;
        if (!bwi_inner(code, d)) return ms_fail(MS_ERR_INVALID, "format code %u of block %llu", code,
                                                (unsigned long long)b);
        out.codes[b] = (uint8_t)code;
        const uint64_t q = ms_align64(pos + 1);
        if (q > size_compr) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: view past sizeCompr");
        uint64_t sz = 0;
        if (code == 0) sz = 8 * B;
        else if (code <= 64) sz = code * B / 8;
        else if (code == MS_BWI_RLE) {
            uint64_t done = 0;
            while (done < B) {
                if (q + sz + 16 > size_compr) return ms_fail(MS_ERR_INVALID, "rle view past sizeCompr");
                uint64_t len;
                memcpy(&len, img + q + sz + 8, 8);
                if (len == 0 || len > B - done) return ms_fail(MS_ERR_INVALID, "rle view: run length %llu",
                                                               (unsigned long long)len);
                done += len;
                sz += 16;
            }
        } else {
            std::vector<uint64_t> ix, rf;
            MS_TRY(walk(d, img + q, B, size_compr - q, ix, rf, &sz));
            const int k = kind_of(code);
            if (k < 3) {
                for (uint64_t e : ix) idx[k].push_back(e + (q << 8));
                if (k == BWK_FOR) refs[k].insert(refs[k].end(), rf.begin(), rf.end());
            }
        }
        if (q + sz > size_compr) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: view past sizeCompr");
        off[b] = q;
        size[b] = sz;
        pos = q + sz;
    }
    bwi_layout(out.h, out.codes.data(), out.lists);
    for (uint64_t b = 0; b < nb; b++) {
        out.h.v[b].off = off[b];
        out.h.v[b].size = size[b];
    }
    out.h.C = nb ? pos : 0;
    out.idx.clear();
    out.refs.assign(bwi_refs_entries(out.h), 0);
    for (int k = 0; k < 3; k++) out.idx.insert(out.idx.end(), idx[k].begin(), idx[k].end());
    std::copy(refs[BWK_FOR].begin(), refs[BWK_FOR].end(), out.refs.begin() + out.h.refs_off[BWK_FOR]);
    if (out.idx.size() != bwi_index_entries(out.h)) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: index");
    return MS_OK;
}

int bwi_finish_upload(ms_col *c, bwi_host &hb, temps &sc)
{
    hipStream_t s = sc.s;
    c->bwi = new (std::nothrow) ms_bwi(hb.h);
    if (!c->bwi) return ms_fail(MS_ERR_NOMEM, "host allocation");
    MS_TRY(upload_tab(c, hb.h, hb.codes, hb.lists, s));
    // delta carries: a decoding pass per delta view
    for (uint64_t j = 0; j < hb.h.lcount[BWK_DELTA]; j++) {
        const ms_col g = group_col(c, BWK_DELTA, (int64_t)j);
        MS_TRY(launch_delta_carries(&g, sc));
    }
    return MS_OK;
}

}  // namespace ms
