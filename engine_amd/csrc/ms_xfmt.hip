// ms_xfmt.hip -- the reference's remaining candidate formats on gfx950:
//   * static_vbp_f<vbp_padding_l<bw, S>> (vbp_padding.h:52-205): a block
//     is 64/bw codes per lane; lane j's word = OR_k v[k S + j] << k bw, no
//     code crosses a word;
//   * k_wise_ns_f<2> (k_wise_ns.h:43-262): per pair of values a descriptor
//     byte (leading zero bytes of each, 3 bits each) and the low bytes;
//   * group_simple_f<E, uint64_t, 8E> (group_simple.h:64-365): groups of E
//     values, a greedy choice of one of 15 (groups, width) modes per
//     8E-byte block, 4-bit selectors in front of the blocks.
// Column-level morph (the compressed part of floor(n / bs) * bs values as
// one morph_batch call, the rest raw behind the gap) and decompression;
// the operators read these columns through decompression.
//
// Derived block index (not part of the image):
//   k_wise_ns: per 32 pairs the byte offset of their first descriptor;
//   group_simple: per block (data byte offset << 8) | code, code = the
//     selector (0..14) or 16 + R for an incomplete block of R groups, and
//     in refs the block's first value.
// group_simple's greedy parse is a chain of steps (each takes 1..64 groups
// and depends only on the groups ahead of it): every group position's step
// is computed in parallel, then each segment of SEG groups is walked from
// each of the 64 positions a chain can enter it at (one lane each), the
// segments' entry points follow from composing those maps, and a last walk
// writes the selectors.
#include "ms_units.hpp"

namespace ms {

// ------------------------------------------------------------------ common
// Column meta once the compressed part ends at *d_c (device): the rest of
// pl_m - n_c values goes raw behind the gap.
__global__ void k_x_finish(ms_dev_meta *meta, const uint64_t *d_m, const uint64_t *d_c, uint64_t bs, uint64_t gap,
                           uint64_t cap, uint64_t *d_rest_off)
{
    const uint64_t m = *d_m, nc = m / bs * bs, C = *d_c;
    const uint64_t R = ms_align64(C + gap);
    const uint64_t rest = m - nc;
    meta->count = m;
    meta->size_compr = C;
    meta->size_used = rest ? R + 8 * rest : C;
    meta->aux[0] = rest ? R + 8 * rest : R;
    meta->aux[1] = R;
    *d_rest_off = R;
    if (R + 8 * rest > cap) atomicCAS((unsigned long long *)&meta->status, 0ull, (unsigned long long)MS_ERR_CAPACITY);
}

__global__ __launch_bounds__(WG) void k_x_rest(const uint64_t *v, const uint64_t *d_m, uint64_t bs,
                                               const uint64_t *d_rest_off, uint64_t cap, uint8_t *out)
{
    const uint64_t m = *d_m, nc = m / bs * bs, rest = m - nc, R = *d_rest_off;
    if (R + 8 * rest > cap) return;   // reported by k_x_finish
    uint64_t *dst = (uint64_t *)(out + R);
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < rest; i += (uint64_t)gridDim.x * WG)
        dst[i] = v[nc + i];
}

static uint32_t grid_n(uint64_t n) { return grid_for(n, WG * 4ull, 4096); }

// ------------------------------------------------------------------ vbp_padding
__global__ __launch_bounds__(WG) void k_pad_pack(const uint64_t *v, const uint64_t *d_m, uint32_t bw, uint32_t S,
                                                 uint64_t *out, uint64_t *d_c)
{
    const uint32_t per = 64 / bw;
    const uint64_t bs = (uint64_t)per * S;
    const uint64_t nblk = *d_m / bs, words = nblk * S;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < words; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = i / S, j = i % S;
        const uint64_t *x = v + b * bs + j;
        uint64_t w = 0;
        for (uint32_t k = 0; k < per; k++) w |= x[(uint64_t)k * S] << (k * bw);   // unmasked, as the reference
        out[i] = w;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_c = words * 8;
}

__global__ __launch_bounds__(WG) void k_pad_unpack(ms_colv c, uint64_t *out)
{
    const ms_parts pt = parts_of(c);
    const uint32_t bw = c.g.bw, S = c.g.lanes, per = 64 / bw;
    const uint64_t bs = (uint64_t)per * S, words = pt.n_compr / bs * S;
    const uint64_t mask = bw >= 64 ? ~0ull : (1ull << bw) - 1ull;
    const uint64_t *in = (const uint64_t *)c.data;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < words; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = i / S, j = i % S, w = in[i];
        uint64_t *o = out + b * bs + j;
        for (uint32_t k = 0; k < per; k++) o[(uint64_t)k * S] = (w >> (k * bw)) & mask;
    }
}

// ------------------------------------------------------------------ k_wise_ns
__device__ __forceinline__ uint32_t zero_bytes(uint64_t x) { return (uint32_t)__builtin_clzll(x | 1ull) >> 3; }

__global__ __launch_bounds__(WG) void k_kwns_sizes(const uint64_t *v, const uint64_t *d_m, uint64_t pairs_bound,
                                                   uint64_t *sz)
{
    const uint64_t pairs = *d_m / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pairs_bound; i += (uint64_t)gridDim.x * WG)
        sz[i] = i < pairs ? 17 - zero_bytes(v[2 * i]) - zero_bytes(v[2 * i + 1]) : 0;
}

__global__ __launch_bounds__(WG) void k_kwns_write(const uint64_t *v, const uint64_t *d_m, const uint64_t *off,
                                                   uint8_t *out, uint64_t *index, uint64_t *d_c)
{
    const uint64_t pairs = *d_m / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pairs; i += (uint64_t)gridDim.x * WG) {
        const uint64_t x0 = v[2 * i], x1 = v[2 * i + 1];
        const uint32_t z0 = zero_bytes(x0), z1 = zero_bytes(x1);
        uint8_t *o = out + off[i];
        *o++ = (uint8_t)((z1 << 3) | z0);
        for (uint32_t b = 0; b < 8 - z0; b++) *o++ = (uint8_t)(x0 >> (8 * b));
        for (uint32_t b = 0; b < 8 - z1; b++) *o++ = (uint8_t)(x1 >> (8 * b));
        if (i % (MS_KWNS_PER_ENTRY / 2) == 0) index[i / (MS_KWNS_PER_ENTRY / 2)] = off[i];
        if (i == pairs - 1) *d_c = off[i] + 17 - z0 - z1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && pairs == 0) *d_c = 0;
}

// one thread walks the 32 pairs of an index entry
__global__ __launch_bounds__(WG) void k_kwns_unpack(ms_colv c, uint64_t *out)
{
    const ms_parts pt = parts_of(c);
    const uint64_t pairs = pt.n_compr / 2, per = MS_KWNS_PER_ENTRY / 2;
    const uint64_t nent = (pairs + per - 1) / per;
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < nent; e += (uint64_t)gridDim.x * WG) {
        const uint8_t *p = c.data + gload(c.index + e);
        const uint64_t i1 = (e + 1) * per < pairs ? (e + 1) * per : pairs;
        for (uint64_t i = e * per; i < i1; i++) {
            const uint32_t d = *p++, z0 = d & 7, z1 = (d >> 3) & 7;
            uint64_t x0 = 0, x1 = 0;
            for (uint32_t b = 0; b < 8 - z0; b++) x0 |= (uint64_t)*p++ << (8 * b);
            for (uint32_t b = 0; b < 8 - z1; b++) x1 |= (uint64_t)*p++ << (8 * b);
            out[2 * i] = x0;
            out[2 * i + 1] = x1;
        }
    }
}

// ------------------------------------------------------------------ group_simple
__constant__ uint8_t c_gs_num[15] = {64, 32, 21, 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1};
__constant__ uint8_t c_gs_bits[15] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 21, 32, 64};
static constexpr uint8_t GS_STOP = 0xff;
static constexpr uint64_t GS_SEG = 4096;   // groups per segment of the chain walk

// The smallest mode whose width holds the group's OR (group_simple.h:282-285).
__device__ __forceinline__ uint8_t gs_min_mode(uint64_t o)
{
    const uint32_t b = 64 - (uint32_t)__builtin_clzll(o | 1ull);   // effective bits
    uint8_t i = 0;
    while (c_gs_bits[i] < b) i++;
    return i;
}

__global__ __launch_bounds__(WG) void k_gs_modes(const uint64_t *v, uint64_t G, uint32_t E, uint8_t *mode)
{
    for (uint64_t g = (uint64_t)blockIdx.x * WG + threadIdx.x; g < G; g += (uint64_t)gridDim.x * WG) {
        uint64_t o = 0;
        for (uint32_t j = 0; j < E; j++) o |= v[g * E + j];
        mode[g] = gs_min_mode(o);
    }
}

// The selector a block starting at group p takes, or GS_STOP when the
// greedy choice there does not get its n groups (group_simple.h:286-312).
__global__ __launch_bounds__(WG) void k_gs_steps(const uint8_t *mode, uint64_t G, uint8_t *step)
{
    for (uint64_t p = (uint64_t)blockIdx.x * WG + threadIdx.x; p < G; p += (uint64_t)gridDim.x * WG) {
        const uint64_t avail = G - p < 64 ? G - p : 64;
        uint8_t res = GS_STOP;
        for (uint32_t i = mode[p]; i < 15; i++) {
            const uint64_t n = c_gs_num[i], maxpos = n < avail ? n : avail;
            uint64_t off = 1;   // mode[p] <= i
            while (off < maxpos && mode[p + off] <= i) off++;
            if (off == maxpos) {
                res = maxpos == n ? (uint8_t)i : GS_STOP;
                break;
            }
        }
        step[p] = res;
    }
}

// Segment s covers groups [s SEG, (s+1) SEG).  Lane e walks the chain from
// group s SEG + e: how many blocks it takes in the segment, and where it
// leaves (its first position past the segment minus the segment end,
// 0..63) or, when it stops inside, 64 + the stop offset in the segment.
__global__ __launch_bounds__(64) void k_gs_seg_maps(const uint8_t *step, uint64_t G, uint64_t nseg, uint32_t *exit_,
                                                    uint32_t *cnt)
{
    const uint32_t e = threadIdx.x;
    for (uint64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        const uint64_t s0 = s * GS_SEG, s1 = s0 + GS_SEG < G ? s0 + GS_SEG : G;
        uint64_t p = s0 + e;
        uint32_t k = 0, ex;
        if (p >= s1) {
            ex = (uint32_t)(p - s1);   // entry past the segment: passes it untouched
        } else {
            for (;;) {
                const uint8_t st = step[p];
                if (st == GS_STOP) {
                    ex = 64 + (uint32_t)(p - s0);
                    break;
                }
                k++;
                p += c_gs_num[st];
                if (p >= s1) {
                    ex = (uint32_t)(p - s1);
                    break;
                }
            }
        }
        exit_[s * 64 + e] = ex;
        cnt[s * 64 + e] = k;
    }
}

// Follows the chain from group 0 through the segments: each segment's
// entry offset and first selector number, the selector count and the
// first group of the incomplete block.
__global__ void k_gs_compose(const uint32_t *exit_, const uint32_t *cnt, uint64_t nseg, uint64_t G, uint32_t *entry,
                             uint64_t *sel_base, uint64_t *d_cs, uint64_t *d_stop)
{
    uint32_t e = 0;
    uint64_t base = 0, stop = G;
    for (uint64_t s = 0; s < nseg; s++) {
        entry[s] = e;
        sel_base[s] = base;
        const uint32_t x = exit_[s * 64 + e];
        base += cnt[s * 64 + e];
        if (x >= 64) {
            stop = s * GS_SEG + (x - 64);
            for (uint64_t t = s + 1; t < nseg; t++) { entry[t] = 0xffffffffu; sel_base[t] = base; }
            break;
        }
        e = x;
    }
    *d_cs = base;
    *d_stop = stop;
}

// The chain's walk through each segment again, writing the selectors and
// the first group of every block.
__global__ void k_gs_emit_sels(const uint8_t *step, uint64_t G, uint64_t nseg, const uint32_t *entry,
                               const uint64_t *sel_base, uint8_t *sel, uint64_t *first)
{
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (uint64_t)gridDim.x * blockDim.x) {
        if (entry[s] == 0xffffffffu) continue;
        const uint64_t s0 = s * GS_SEG, s1 = s0 + GS_SEG < G ? s0 + GS_SEG : G;
        uint64_t p = s0 + entry[s], k = sel_base[s];
        while (p < s1) {
            const uint8_t st = step[p];
            if (st == GS_STOP) break;
            sel[k] = st;
            first[k] = p;
            k++;
            p += c_gs_num[st];
        }
    }
}

// Header, selector bytes (low nibble first) and the data area's offset:
// aligned to 8E bytes from the image base (the call starts at base0).
__global__ __launch_bounds__(WG) void k_gs_header(const uint8_t *sel, const uint64_t *d_cs, const uint64_t *d_stop,
                                                  uint64_t G, uint32_t E, uint64_t base0, uint8_t *out,
                                                  uint64_t *d_data)
{
    const uint64_t cs = *d_cs, nbytes = (cs + 1) / 2;
    for (uint64_t b = (uint64_t)blockIdx.x * WG + threadIdx.x; b < nbytes; b += (uint64_t)gridDim.x * WG)
        out[9 + b] = (uint8_t)(sel[2 * b] | (2 * b + 1 < cs ? sel[2 * b + 1] << 4 : 0));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int q = 0; q < 8; q++) out[q] = (uint8_t)(cs >> (8 * q));
        out[8] = (uint8_t)(G - *d_stop);
        const uint64_t a = 8ull * E;
        const uint64_t data = (base0 + 9 + nbytes + a - 1) / a * a - base0;
        for (uint64_t q = 9 + nbytes; q < data; q++) out[q] = 0;   // the alignment padding reads as zeros
        *d_data = data;
    }
}

// Block k of the chain (and the incomplete block as k = cs): lane j of a
// thread packs its E-lane word; index entry and first value per block.
__global__ __launch_bounds__(WG) void k_gs_pack(const uint64_t *v, const uint8_t *sel, const uint64_t *first,
                                                const uint64_t *d_cs, const uint64_t *d_stop, const uint64_t *d_data,
                                                uint64_t G, uint32_t E, uint64_t blk_bound, uint8_t *out,
                                                uint64_t *index, uint64_t *refs, uint64_t *d_c, uint64_t *d_nblk)
{
    const uint64_t cs = *d_cs, stop = *d_stop, data = *d_data;
    const uint64_t R = G - stop, nblk = cs + (R ? 1 : 0);
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nblk * E && i < blk_bound * E;
         i += (uint64_t)gridDim.x * WG) {
        const uint64_t k = i / E, j = i % E;
        uint64_t g0, n, b, code;
        if (k < cs) {
            g0 = first[k];
            n = c_gs_num[sel[k]];
            b = c_gs_bits[sel[k]];
            code = sel[k];
        } else {
            g0 = stop;
            n = R;
            b = 64 / R;
            code = 16 + R;
        }
        uint64_t w = 0;
        for (uint64_t q = 0; q < n; q++) w |= v[(g0 + q) * E + j] << (q * b);
        const uint64_t off = data + 8ull * E * k;
        *(uint64_t *)(out + off + 8 * j) = w;
        if (j == 0) {
            index[k] = (off << 8) | code;
            refs[k] = g0 * E;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *d_c = data + 8ull * E * nblk;
        *d_nblk = nblk;
    }
}

// Decompression: lane j of block k writes its n values (the column's
// block count is in its meta's aux[2]).
__global__ __launch_bounds__(WG) void k_gs_unpack(ms_colv c, uint64_t *out)
{
    const uint32_t E = c.g.lanes;
    const uint64_t nblk = c.meta->aux[2];
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nblk * E; i += (uint64_t)gridDim.x * WG) {
        const uint64_t k = i / E, j = i % E;
        const uint64_t e = gload(c.index + k), first = gload(c.refs + k);
        const uint32_t code = (uint32_t)(e & 0xff);
        uint32_t n, b;
        bool last_masked;
        if (code < 15) {
            n = c_gs_num[code];
            b = c_gs_bits[code];
            last_masked = false;   // the generated unpack leaves the last group unmasked
        } else {
            n = code - 16;
            b = 64 / n;
            last_masked = true;
        }
        const uint64_t w = gload((const uint64_t *)(c.data + (e >> 8)) + j);
        const uint64_t mask = b >= 64 ? ~0ull : (1ull << b) - 1ull;
        for (uint32_t q = 0; q < n; q++) {
            const uint64_t x = w >> (q * b);
            out[first + (uint64_t)q * E + j] = (q + 1 < n || last_masked) ? x & mask : x;
        }
    }
}

// ------------------------------------------------------------------ launchers
int launch_xfmt_compress(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, uint64_t base0,
                         temps &sc)
{
    hipStream_t s = sc.s;
    const ms_geom &g = out->g;
    uint64_t *d = (uint64_t *)sc.get(8 * 8);
    if (!d) return ms_fail(MS_ERR_NOMEM, "compress scratch");
    uint64_t *d_c = d, *d_rest = d + 1;
    switch (g.kind) {
    case MS_STATIC_VBP_PAD:
        hipLaunchKernelGGL(k_pad_pack, dim3(grid_n(m_bound / g.bs * g.lanes + 1)), dim3(WG), 0, s, d_v, d_m, g.bw,
                           g.lanes, (uint64_t *)out->data, d_c);
        break;
    case MS_K_WISE_NS: {
        const uint64_t pb = m_bound / 2 + 1;
        uint64_t *sz = (uint64_t *)sc.get(8 * pb + 64), *off = (uint64_t *)sc.get(8 * pb + 64);
        if (!sz || !off) return ms_fail(MS_ERR_NOMEM, "k_wise_ns scratch");
        if (out->index_cap < m_bound / MS_KWNS_PER_ENTRY + 1) return ms_fail(MS_ERR_CAPACITY, "k_wise_ns index");
        hipLaunchKernelGGL(k_kwns_sizes, dim3(grid_n(pb)), dim3(WG), 0, s, d_v, d_m, pb, sz);
        MS_TRY(launch_exclusive_scan(sz, off, pb, sc));
        hipLaunchKernelGGL(k_kwns_write, dim3(grid_n(pb)), dim3(WG), 0, s, d_v, d_m, (const uint64_t *)off,
                           out->data, out->index, d_c);
        break;
    }
    case MS_GROUP_SIMPLE: {
        // the compressed part has floor(m / E) groups; the host knows a bound
        const uint32_t E = g.lanes;
        const uint64_t Gb = m_bound / E;
        // the kernels below need the exact group count: read it back
        uint64_t m = 0;
        MS_HIP(hipMemcpyAsync(&m, d_m, 8, hipMemcpyDeviceToHost, s));
        MS_HIP(hipStreamSynchronize(s));
        const uint64_t G = m / E;
        if (G > Gb) return ms_fail(MS_ERR_CAPACITY, "group count above its bound");
        const uint64_t nseg = (G + GS_SEG - 1) / GS_SEG;
        uint8_t *mode = (uint8_t *)sc.get(G + 64), *step = (uint8_t *)sc.get(G + 64), *sel = (uint8_t *)sc.get(G + 64);
        uint32_t *ex = (uint32_t *)sc.get(4 * 64 * nseg + 64), *cnt = (uint32_t *)sc.get(4 * 64 * nseg + 64);
        uint32_t *entry = (uint32_t *)sc.get(4 * nseg + 64);
        uint64_t *sel_base = (uint64_t *)sc.get(8 * nseg + 64), *first = (uint64_t *)sc.get(8 * G + 64);
        if (!mode || !step || !sel || !ex || !cnt || !entry || !sel_base || !first)
            return ms_fail(MS_ERR_NOMEM, "group_simple scratch");
        if (out->index_cap < G + 2 || out->refs_cap < G + 2) return ms_fail(MS_ERR_CAPACITY, "group_simple index");
        uint64_t *d_cs = d + 2, *d_stop = d + 3, *d_data = d + 4;
        if (G == 0) {   // no morph_batch call at all (morph.h:182-185): an empty compressed part
            MS_TRY(launch_fill_u64(d_c, 1, 0, s));
            MS_TRY(launch_fill_u64(&out->meta->aux[2], 1, 0, s));
            break;
        }
        {
            hipLaunchKernelGGL(k_gs_modes, dim3(grid_n(G)), dim3(WG), 0, s, d_v, G, E, mode);
            hipLaunchKernelGGL(k_gs_steps, dim3(grid_n(G)), dim3(WG), 0, s, (const uint8_t *)mode, G, step);
            hipLaunchKernelGGL(k_gs_seg_maps, dim3(grid_for(nseg, 1, 65536)), dim3(64), 0, s, (const uint8_t *)step,
                               G, nseg, ex, cnt);
        }
        hipLaunchKernelGGL(k_gs_compose, dim3(1), dim3(1), 0, s, (const uint32_t *)ex, (const uint32_t *)cnt, nseg,
                           G, entry, sel_base, d_cs, d_stop);
        hipLaunchKernelGGL(k_gs_emit_sels, dim3(grid_for(nseg, 64, 4096)), dim3(64), 0, s, (const uint8_t *)step,
                               G, nseg, (const uint32_t *)entry, (const uint64_t *)sel_base, sel, first);
        hipLaunchKernelGGL(k_gs_header, dim3(grid_n(G / 2 + 1)), dim3(WG), 0, s, (const uint8_t *)sel,
                           (const uint64_t *)d_cs, (const uint64_t *)d_stop, G, E, base0, out->data, d_data);
        hipLaunchKernelGGL(k_gs_pack, dim3(grid_n((G + 1) * E)), dim3(WG), 0, s, d_v, (const uint8_t *)sel,
                           (const uint64_t *)first, (const uint64_t *)d_cs, (const uint64_t *)d_stop,
                           (const uint64_t *)d_data, G, E, G + 1, out->data, out->index, out->refs, d_c,
                           &out->meta->aux[2]);
        break;
    }
    default:
        return ms_fail(MS_ERR_INVALID, "no compressor for format kind %u", g.kind);
    }
    MS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_x_finish, dim3(1), dim3(1), 0, s, out->meta, d_m, (const uint64_t *)d_c, g.bs, g.gap,
                       out->size_alloc, d_rest);
    hipLaunchKernelGGL(k_x_rest, dim3(grid_for(g.bs + 1, WG, 64)), dim3(WG), 0, s, d_v, d_m, g.bs,
                       (const uint64_t *)d_rest, out->size_alloc, out->data);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

int launch_xfmt_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s)
{
    ms_colv v;
    MS_TRY(colv_of(in, v));
    const uint64_t nb = in->count_bound;
    switch (v.g.kind) {
    case MS_STATIC_VBP_PAD:
        hipLaunchKernelGGL(k_pad_unpack, dim3(grid_n(nb / v.g.bs * v.g.lanes + 1)), dim3(WG), 0, s, v, d_out);
        break;
    case MS_K_WISE_NS:
        hipLaunchKernelGGL(k_kwns_unpack, dim3(grid_for(nb / MS_KWNS_PER_ENTRY + 1, WG, 4096)), dim3(WG), 0, s, v,
                           d_out);
        break;
    case MS_GROUP_SIMPLE:
        hipLaunchKernelGGL(k_gs_unpack, dim3(grid_n(in->index_cap * v.g.lanes)), dim3(WG), 0, s, v, d_out);
        break;
    default:
        return ms_fail(MS_ERR_INVALID, "no decompressor for format kind %u", v.g.kind);
    }
    MS_HIP(hipGetLastError());
    return MS_OK;
}

}  // namespace ms
