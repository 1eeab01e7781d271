// ms_decode.hip -- read side of the compression-aware path on gfx950:
//   * fused decompress + select + agg_sum (decompress_and_process_batch with
//     the select core feeding the agg_sum core, format.h:164-195,
//     select_compr.h:80-89, agg_sum_compr.h:57-71) for every format,
//   * decompression (morph_t<ve, uncompr_f, F>, morph.h:203-241),
//   * select's position list (select_compr.h:100-260) as count + compact.
//
// Mapping: a wave decodes a 512-value unit.  For a vertical-bit-packed
// layout with S lanes, thread t handles lane j = t % S and the 8 codes
// k = 8g .. 8g+7 with g = unit*(64/S) + t/S (values S*k + j).
#include <map>
#include <mutex>


#include "ms_units.hpp"

namespace ms {

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s)
{
    if (n == 0) return MS_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, WG * 4ull, 2048)), dim3(WG), 0, s, p, n, v);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// One wave decodes one 512-value unit per iteration (grid-stride); the
// next unit's index entry is fetched one iteration ahead.  Unit -> block
// and block -> for_f block are shifts (the host only takes this path when
// the ratios are powers of two).  dynamic_vbp takes its per-bit-width
// predicate from a table built on the host; for_f specialises it per block
// around the block's reference value.
template <bool IS_FOR, bool LEFORM>
__global__ __launch_bounds__(WG) void k_select_sum_blocks(ms_colv c, ms_pred p, ms_bwtab tab, uint64_t *res)
{
    const ms_parts pt = parts_of(c);
    const uint32_t lg_spb = c.g.lg_spb;
    const uint32_t lg_ipo = c.g.lg_ipo;
    const uint64_t n = pt.n_compr >> 9;
    const uint64_t nq = (n + 3) >> 2;   // groups of 4 units
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & 7, gl = t >> 3;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    uint64_t acc = 0;
    uint64_t q = (uint64_t)blockIdx.x * WAVES + wave_index();
    // lanes 0..3 hold the group's index entries (and for_f references),
    // requested one group ahead
    uint64_t my_e = 0, my_r = 0;
    if (q < nq && t < 4 && q * 4 + t < n) {
        my_e = gload(c.index + ((q * 4 + t) >> lg_spb));
        if (IS_FOR) my_r = gload(c.refs + (((q * 4 + t) >> lg_spb) >> lg_ipo));
    }
    for (; q < nq; q += W) {
        const uint64_t u0 = q * 4;
        const uint32_t cnt = n - u0 < 4 ? (uint32_t)(n - u0) : 4u;
        uint64_t e[4], r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            e[u] = readlane64(my_e, u);
            r[u] = IS_FOR ? readlane64(my_r, u) : 0ull;
        }
        const uint64_t un = u0 + 4 * W + t;
        my_e = my_r = 0;
        if (q + W < nq && t < 4 && un < n) {
            my_e = gload(c.index + (un >> lg_spb));
            if (IS_FOR) my_r = gload(c.refs + ((un >> lg_spb) >> lg_ipo));
        }
        vbp_units_ss<IS_FOR, LEFORM>(c.data, e, r, cnt, u0, mask, p, tab, j, gl, acc);
    }
    wg_partial(acc, res);
}

// static_vbp_f<vbp_l<BW, S>> with 64 % S == 0.
template <int BW>
__global__ __launch_bounds__(WG) void k_select_sum_static(ms_colv c, ms_pred p, uint64_t *res)
{
    const ms_parts pt = parts_of(c);
    const uint32_t S = c.g.lanes, lgS = c.g.lg_lanes;
    const uint64_t gpu = 64 >> lgS;
    const uint64_t ngroups = pt.n_compr >> (lgS + 3);
    const uint64_t nunits = (ngroups + gpu - 1) >> (6 - lgS);
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & (S - 1);
    const uint64_t maxcode = BW >= 64 ? ~0ull : (1ull << BW) - 1ull;
    const ms_cpred cp = code_pred(p, 0, maxcode);
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    uint64_t acc = 0;
    if (cp.mode != PM_NONE) {
        const uint64_t *base = (const uint64_t *)c.data;
        uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index();
        // four units a step, every unit's words requested before the first is cut
        constexpr int NW = vbp_traits<BW>::NW;
        for (; u + 3 * nwaves < nunits; u += 4 * nwaves) {
            uint64_t w0[NW], w1[NW], w2[NW], w3[NW];   // named arrays: they stay in registers
            uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            const uint64_t g0 = u * gpu + (t >> lgS), g1 = g0 + nwaves * gpu, g2 = g1 + nwaves * gpu,
                           g3 = g2 + nwaves * gpu;
            if (g0 < ngroups) vbp_load<BW>(base, S, j, g0, w0, s0);
            if (g1 < ngroups) vbp_load<BW>(base, S, j, g1, w1, s1);
            if (g2 < ngroups) vbp_load<BW>(base, S, j, g2, w2, s2);
            if (g3 < ngroups) vbp_load<BW>(base, S, j, g3, w3, s3);
            if (g0 < ngroups) ss_loaded<BW>(w0, s0, cp, 0, acc);
            if (g1 < ngroups) ss_loaded<BW>(w1, s1, cp, 0, acc);
            if (g2 < ngroups) ss_loaded<BW>(w2, s2, cp, 0, acc);
            if (g3 < ngroups) ss_loaded<BW>(w3, s3, cp, 0, acc);
        }
        for (; u < nunits; u += nwaves) {
            const uint64_t g = u * gpu + (t >> lgS);
            if (g < ngroups) ss_unit<BW>(base, S, j, g, cp, 0, acc);
        }
    }
    wg_partial(acc, res);
}

// Uncompressed values: the whole of an uncompr_f column, or the rest part.
__global__ __launch_bounds__(WG) void k_select_sum_uncompr(ms_colv c, ms_pred p, uint64_t *res)
{
    const ms_parts pt = parts_of(c);
    const uint64_t *x = (const uint64_t *)(c.data + pt.rest_off);
    const uint64_t n = pt.n_rest;
    uint64_t acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * WG;
    uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x;
    for (; 2 * i + 1 < n; i += stride) {
        const u64x2 v = __builtin_nontemporal_load((const u64x2 *)x + i);
        acc += pred_eval(p, v.x) ? v.x : 0;
        acc += pred_eval(p, v.y) ? v.y : 0;
    }
    if (2 * i < n && 2 * i + 1 == n) {
        const uint64_t v = x[2 * i];
        acc += pred_eval(p, v) ? v : 0;
    }
    wg_partial(acc, res);
}

// *d_n decompressed values.
__global__ __launch_bounds__(WG) void k_select_sum_values(const uint64_t *x, const uint64_t *d_n, ms_pred p,
                                                          uint64_t *res)
{
    const uint64_t n = *d_n;
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        acc += pred_eval(p, x[i]) ? x[i] : 0ull;
    wg_partial(acc, res);
}

// Any layout readable by read_compr, one value per thread.
__global__ __launch_bounds__(WG) void k_select_sum_generic(ms_colv c, ms_pred p, uint64_t *res)
{
    const ms_parts pt = parts_of(c);
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pt.n_compr; i += (uint64_t)gridDim.x * WG) {
        const uint64_t x = read_compr(c, i);
        if (pred_eval(p, x)) acc += x;
    }
    wg_partial(acc, res);
}

// rle_f: a run of `len` copies of v contributes len * v when v passes.
static constexpr uint64_t RLE_CHUNK = 512;   // runs per wave and step

__global__ __launch_bounds__(WG) void k_select_sum_rle(ms_colv c, ms_pred p, uint64_t *res)
{
    const uint64_t runs = c.meta->size_used / 16;
    const u64x2 *w = (const u64x2 *)c.data;
    const uint64_t nch = (runs + RLE_CHUNK - 1) / RLE_CHUNK;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    const uint32_t t = threadIdx.x & 63;
    uint64_t acc = 0;
    for (uint64_t ch = (uint64_t)blockIdx.x * WAVES + wave_index(); ch < nch; ch += W) {
        const uint64_t r0 = ch * RLE_CHUNK, r1 = r0 + RLE_CHUNK < runs ? r0 + RLE_CHUNK : runs;
        acc += rle_runs_ss(w, r0, r1, p, t);
    }
    wg_partial(acc, res);
}

// ------------------------------------------------------------------
// delta_f: x = lane-wise running sum of the decoded deltas (delta.h:
// 229-261).  The column's derived index holds, per 2048-value tile, the
// running sums of the S lanes at the tile's start (its "carries"), so every
// tile decodes independently: a workgroup (4 waves x 512 values) scans its
// tile locally and adds the carries.  SRC 0 decodes 8-lane indexed blocks,
// SRC 1 reads deltas already decoded into `arr`.
// ------------------------------------------------------------------
static constexpr uint64_t DELTA_TILE = MS_DELTA_TILE;
static_assert(MS_DELTA_TILE == WAVES * 512, "a delta tile is one workgroup's 4 units");

template <int SRC>
// e_v: unit u's index entry (SRC 0), loaded by the caller.
__device__ __forceinline__ void delta_load(const ms_colv &c, const uint64_t *arr, uint64_t u, uint64_t nunits,
                                           uint32_t S, uint32_t j, uint32_t gl, uint64_t e_v, uint64_t (&d)[8])
{
    if (u >= nunits) {
#pragma unroll
        for (int m = 0; m < 8; m++) d[m] = 0;
        return;
    }
    if (SRC == 0) {
        const uint64_t sub = u & ((1ull << c.g.lg_spb) - 1);
        const uint64_t e = uniform64(e_v);
        dec_dispatch((uint32_t)(e & 0xff), (const uint64_t *)(c.data + (e >> 8)), 8, j, sub * 8 + gl, d);
    } else {
        const uint64_t *a = arr + u * 512;
#pragma unroll
        for (int m = 0; m < 8; m++) d[m] = a[S * (8 * gl + m) + j];
    }
}

// Lane-wise scan of one 2048-value tile: returns, per thread, the
// exclusive prefix of its 8 values inside the tile (d[] becomes the
// inclusive prefix within the thread).  tile_tot receives lane j's total.
__device__ __forceinline__ uint64_t delta_tile_scan(uint64_t (&d)[8], uint32_t S, uint32_t j, uint32_t t,
                                                    uint32_t wv, uint64_t (*s_wtot)[8], uint64_t &tile_tot)
{
#pragma unroll
    for (int m = 1; m < 8; m++) d[m] += d[m - 1];
    const uint64_t tot = d[7];
    uint64_t incl = tot;
    for (uint32_t o = S; o < 64; o <<= 1) {
        const uint64_t v = __shfl_up(incl, o, MS_WAVE);
        if (t >= o) incl += v;
    }
    if (t >= 64 - S) s_wtot[wv][j] = incl;
    __syncthreads();
    uint64_t wpre = 0;
    tile_tot = 0;
    for (uint32_t w = 0; w < WAVES; w++) {
        const uint64_t v = s_wtot[w][j];
        if (w < wv) wpre += v;
        tile_tot += v;
    }
    __syncthreads();
    return wpre + incl - tot;
}

// MODE 0 writes x, MODE 1 folds select + sum into per-workgroup partials.
template <int MODE, int SRC>
__global__ __launch_bounds__(WG) void k_delta_tiles(ms_colv c, const uint64_t *arr, uint64_t *out, ms_pred p,
                                                    uint64_t *part)
{
    __shared__ uint64_t s_wtot[WAVES][8];
    const ms_parts pt = parts_of(c);
    const uint32_t S = c.g.dstep;
    const uint64_t nunits = pt.n_compr >> 9;
    const uint64_t ntiles = (nunits + WAVES - 1) / WAVES;
    const uint32_t wv = wave_index();
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & (S - 1), gl = t >> c.g.lg_dstep;
    uint64_t acc = 0;
    // SRC 0: each unit's index entry is requested one tile ahead
    uint64_t e_n = 0;
    if (SRC == 0 && blockIdx.x * WAVES + wv < nunits) e_n = gload(c.index + ((blockIdx.x * WAVES + wv) >> c.g.lg_spb));
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t u = tile * WAVES + wv;
        const uint64_t carry = gload(c.refs + tile * 8 + j);   // needed after the scan: requested first
        const uint64_t e_v = e_n;
        const uint64_t un = (tile + gridDim.x) * WAVES + wv;
        if (SRC == 0) e_n = un < nunits ? gload(c.index + (un >> c.g.lg_spb)) : 0ull;
        uint64_t d[8];
        delta_load<SRC>(c, arr, u, nunits, S, j, gl, e_v, d);
        uint64_t tile_tot;
        const uint64_t base = carry + delta_tile_scan(d, S, j, t, wv, s_wtot, tile_tot);
        if (u < nunits) {
            if (MODE == 0) {
                uint64_t *o = out + u * 512;
#pragma unroll
                for (int m = 0; m < 8; m++) o[S * (8 * gl + m) + j] = base + d[m];
            } else {
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    const uint64_t x = base + d[m];
                    acc += pred_eval(p, x) ? x : 0;
                }
            }
        }
    }
    if (MODE == 1) wg_partial(acc, part);
}

// Select + sum over a delta_f column with 8 lanes and indexed 8-lane
// inner blocks: one wave per tile, no workgroup synchronisation.
__global__ __launch_bounds__(WG) void k_delta_select_sum(ms_colv c, ms_pred p, uint64_t *part)
{
    const ms_parts pt = parts_of(c);
    const uint64_t nunits = pt.n_compr >> 9;
    const uint64_t ntiles = (nunits + 3) >> 2;
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t >> 3, gl = t & 7;   // delta lane j in 8 consecutive threads
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    uint64_t acc = 0;
    uint64_t tile = (uint64_t)blockIdx.x * WAVES + wave_index();
    // each tile's index entries are requested one tile ahead
    uint64_t my_e = tile < ntiles ? delta_tile_entries(c.index, c.g.lg_spb, tile, nunits, t) : 0ull;
    for (; tile < ntiles; tile += W) {
        const uint64_t e_next =
            tile + W < ntiles ? delta_tile_entries(c.index, c.g.lg_spb, tile + W, nunits, t) : 0ull;
        delta_tile_ss<false>(c.data, my_e, c.refs, c.g.lg_spb, p, tile, nunits, t, j, gl, acc);
        my_e = e_next;
    }
    wg_partial(acc, part);
}

// Per-tile lane totals, lane-major: sums[j * ntb + tile].
template <int SRC>
__global__ __launch_bounds__(WG) void k_delta_tile_sums(ms_colv c, const uint64_t *arr, uint64_t *sums, uint64_t ntb)
{
    __shared__ uint64_t s_wtot[WAVES][8];
    const ms_parts pt = parts_of(c);
    const uint32_t S = c.g.dstep;
    const uint64_t nunits = pt.n_compr >> 9;
    const uint32_t wv = wave_index();
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & (S - 1), gl = t >> c.g.lg_dstep;
    for (uint64_t tile = blockIdx.x; tile < ntb; tile += gridDim.x) {
        const uint64_t u = tile * WAVES + wv;
        const uint64_t e_v = SRC == 0 && u < nunits ? gload(c.index + (u >> c.g.lg_spb)) : 0ull;
        uint64_t d[8];
        delta_load<SRC>(c, arr, u, nunits, S, j, gl, e_v, d);
        uint64_t tile_tot;
        (void)delta_tile_scan(d, S, j, t, wv, s_wtot, tile_tot);
        if (threadIdx.x < S) sums[(uint64_t)threadIdx.x * ntb + tile] = tile_tot;
    }
}

__global__ __launch_bounds__(WG) void k_delta_carries(const uint64_t *scan, uint64_t ntb, uint32_t S, uint64_t *refs)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < ntb * 8; i += (uint64_t)gridDim.x * WG) {
        const uint64_t tile = i / 8, j = i % 8;
        refs[i] = j < S ? scan[j * ntb + tile] : 0;
    }
}

// ------------------------------------------------------------------
// Decompression kernels.
// ------------------------------------------------------------------
// 8-lane indexed layouts (dynamic_vbp_f, for_f's inner blocks) and
// static_vbp_f<vbp_l<bw, 8>> (LAYOUT 2: unit u's payload is 64 bw bytes at
// u * 64 bw, no index): one resident wave of workgroups, grid-stride over
// groups of 4 units; a group's index entries (and for_f references) are
// loaded lane-parallel one group ahead, and its 4 units' words are all
// requested before the first is decoded.
enum { DL_BLOCKS = 0, DL_FOR = 1, DL_STATIC8 = 2 };

// The end of a decompression kernel: the raw rest behind the compressed
// part (grid-stride), and, when om is given, the output column's whole
// meta (count, 8 bytes a value, status and append slots cleared).
__device__ __forceinline__ void decomp_tail(const ms_colv &c, const ms_parts &pt, uint64_t *out, ms_dev_meta *om)
{
    const uint64_t *rx = (const uint64_t *)(c.data + pt.rest_off);
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pt.n_rest; i += (uint64_t)gridDim.x * WG)
        out[pt.n_compr + i] = gload(rx + i);
    if (om && blockIdx.x == 0 && threadIdx.x == 0) {
        om->count = pt.n;
        om->size_used = 8 * pt.n;
        om->size_compr = 0;
        om->status = 0;
        for (int k = 0; k < 4; k++) om->aux[k] = 0;
    }
}

__global__ void k_set_meta_uncompr(const ms_dev_meta *in, ms_dev_meta *om)
{
    const uint64_t n = in->count;
    om->count = n;
    om->size_used = 8 * n;
    om->size_compr = 0;
    om->status = 0;
    for (int k = 0; k < 4; k++) om->aux[k] = 0;
}

template <int LAYOUT, bool NT>
__global__ __launch_bounds__(WG) void k_decomp_units(ms_colv c, uint64_t *out, ms_dev_meta *om)
{
    const ms_parts pt = parts_of(c);
    const uint32_t lg_spb = LAYOUT == DL_STATIC8 ? 0 : c.g.lg_spb;
    const uint32_t lg_ipo = c.g.lg_ipo;
    const uint64_t n = pt.n_compr >> 9;
    const uint64_t nq = (n + 3) >> 2;
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & 7, gl = t >> 3;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    const uint64_t bw8 = c.g.bw;
    uint64_t q = (uint64_t)blockIdx.x * WAVES + wave_index();
    const auto entry = [&](uint64_t u, uint64_t &e, uint64_t &r) {
        e = r = 0;
        if (LAYOUT == DL_STATIC8) return;
        e = gload(c.index + (u >> lg_spb));
        if (LAYOUT == DL_FOR) r = gload(c.refs + ((u >> lg_spb) >> lg_ipo));
    };
    uint64_t my_e = 0, my_r = 0;
    if (q < nq && t < 4 && q * 4 + t < n) entry(q * 4 + t, my_e, my_r);
    for (; q < nq; q += W) {
        const uint64_t u0 = q * 4;
        const uint32_t cnt = n - u0 < 4 ? (uint32_t)(n - u0) : 4u;
        uint64_t e[4], r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (LAYOUT == DL_STATIC8) {
                e[u] = (((u0 + u) * 64ull * bw8) << 8) | bw8;
                r[u] = 0;
            } else {
                e[u] = readlane64(my_e, u);
                r[u] = LAYOUT == DL_FOR ? readlane64(my_r, u) : 0ull;
            }
        }
        const uint64_t un = u0 + 4 * W + t;
        my_e = my_r = 0;
        if (q + W < nq && t < 4 && un < n) entry(un, my_e, my_r);
        dec_units<NT>(c.data, e, r, cnt, u0, mask, j, gl, out);
    }
    decomp_tail(c, pt, out, om);
}

// delta_f with 8 lanes over indexed 8-lane inner blocks: one wave per
// 2048-value tile (lane-major threads, 8 per delta lane), the tile's index
// entries requested one tile ahead.
template <bool NT>
__global__ __launch_bounds__(WG) void k_delta_decomp(ms_colv c, uint64_t *out, ms_dev_meta *om)
{
    const ms_parts pt = parts_of(c);
    const uint64_t nunits = pt.n_compr >> 9;
    const uint64_t ntiles = (nunits + 3) >> 2;
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t >> 3, gl = t & 7;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    uint64_t tile = (uint64_t)blockIdx.x * WAVES + wave_index();
    uint64_t my_e = tile < ntiles ? delta_tile_entries(c.index, c.g.lg_spb, tile, nunits, t) : 0ull;
    for (; tile < ntiles; tile += W) {
        const uint64_t e_next =
            tile + W < ntiles ? delta_tile_entries(c.index, c.g.lg_spb, tile + W, nunits, t) : 0ull;
        delta_tile_store<NT>(c.data, my_e, c.refs, c.g.lg_spb, tile, nunits, t, j, gl, out);
        my_e = e_next;
    }
    decomp_tail(c, pt, out, om);
}

template <int BW>
__global__ __launch_bounds__(WG) void k_decomp_static(ms_colv c, uint64_t *out)
{
    const ms_parts pt = parts_of(c);
    const uint32_t S = c.g.lanes, lgS = c.g.lg_lanes;
    const uint64_t gpu = 64 >> lgS;
    const uint64_t ngroups = pt.n_compr >> (lgS + 3);
    const uint64_t nunits = (ngroups + gpu - 1) >> (6 - lgS);
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & (S - 1);
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    for (uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index(); u < nunits; u += nwaves) {
        const uint64_t g = u * gpu + (t >> lgS);
        if (g >= ngroups) continue;
        uint64_t d[8];
        dec_unit<BW>((const uint64_t *)c.data, S, j, g, d);
#pragma unroll
        for (int m = 0; m < 8; m++) out[S * (8 * g + m) + j] = d[m];
    }
}

// read_compr per value: the fallback for every other layout (and deltas
// of a delta_f whose inner format is not an 8-lane indexed layout).
__global__ __launch_bounds__(WG) void k_decomp_generic(ms_colv c, uint64_t *out)
{
    const ms_parts pt = parts_of(c);
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pt.n_compr; i += (uint64_t)gridDim.x * WG)
        out[i] = read_compr(c, i);
}

__global__ __launch_bounds__(WG) void k_copy_rest(ms_colv c, uint64_t *out)
{
    const ms_parts pt = parts_of(c);
    const uint64_t *x = (const uint64_t *)(c.data + pt.rest_off);
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < pt.n_rest; i += (uint64_t)gridDim.x * WG)
        out[pt.n_compr + i] = x[i];
}

// rle_f expansion in chunks of RLE_XCH runs: the chunks' value counts
// (k_rle_chunk_lens), their exclusive scan, then one wave per chunk
// (k_rle_expand_chunks): 64 runs at a time are loaded lane-parallel, their
// starts come from a wave scan of the lengths, and the wave writes each
// run's copies with contiguous stores.
static constexpr uint64_t RLE_XCH = 1024;

__device__ __forceinline__ uint64_t block_sum(uint64_t v);

// A workgroup per chunk: each wave sums a quarter of its run lengths.
// *d_nch receives the number of chunks (the scan's length).
__global__ __launch_bounds__(WG) void k_rle_chunk_lens(ms_colv c, uint64_t *clen, uint64_t *d_nch)
{
    static_assert(RLE_XCH % (WAVES * 64) == 0, "a chunk splits into whole wave rows");
    constexpr uint32_t PER = RLE_XCH / (WAVES * 64);   // loads per lane
    const uint64_t runs = c.meta->size_used / 16;
    const uint64_t nch = (runs + RLE_XCH - 1) / RLE_XCH;
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_nch = nch;
    const u64x2 *w = (const u64x2 *)c.data;
    for (uint64_t ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const uint64_t r0 = ch * RLE_XCH + (uint64_t)(threadIdx.x >> 6) * (RLE_XCH / WAVES) + (threadIdx.x & 63);
        uint64_t len[PER];
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) {
            const uint64_t r = r0 + 64 * i;
            len[i] = r < runs ? gload(w + r).y : 0ull;
        }
        uint64_t acc = 0;
#pragma unroll
        for (uint32_t i = 0; i < PER; i++) acc += len[i];
        acc = block_sum(acc);
        if (threadIdx.x == 0) clen[ch] = acc;
    }
}

template <bool NT>
__global__ __launch_bounds__(WG) void k_rle_expand_chunks(ms_colv c, const uint64_t *cstart, uint64_t *out)
{
    const uint64_t runs = c.meta->size_used / 16;
    const uint64_t n = c.meta->count;   // writes stay below it whatever the lengths say
    const uint64_t nch = (runs + RLE_XCH - 1) / RLE_XCH;
    const u64x2 *w = (const u64x2 *)c.data;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    for (uint64_t ch = (uint64_t)blockIdx.x * WAVES + wave_index(); ch < nch; ch += W) {
        uint64_t base = uniform64(gload(cstart + ch));
        const uint64_t r0 = ch * RLE_XCH;
        const uint64_t r1 = r0 + RLE_XCH < runs ? r0 + RLE_XCH : runs;
        u64x2 nx = r0 + t < r1 ? gload(w + r0 + t) : u64x2{0, 0};
        for (uint64_t g = r0; g < r1; g += 64) {
            const u64x2 vl = nx;
            nx = g + 64 + t < r1 ? gload(w + g + 64 + t) : u64x2{0, 0};
            // inclusive scan of the 64 lengths
            uint64_t incl = vl.y;
#pragma unroll
            for (uint32_t o = 1; o < 64; o <<= 1) {
                const uint64_t y = __shfl_up(incl, o, MS_WAVE);
                if (t >= o) incl += y;
            }
            const uint64_t start = base + incl - vl.y;
            base += readlane64(incl, 63);
            const uint32_t nr = r1 - g < 64 ? (uint32_t)(r1 - g) : 64u;
            for (uint32_t k = 0; k < nr; k++) {
                const uint64_t v = readlane64(vl.x, k);
                const uint64_t s0 = readlane64(start, k);
                const uint64_t len = readlane64(vl.y, k);
                const uint64_t e = s0 + len < n && s0 + len >= s0 ? s0 + len : n;
                for (uint64_t i = s0 + t; i < e; i += 64) st64<NT>(out + i, v);
            }
        }
    }
}

// ------------------------------------------------------------------
// Select positions over uncompressed values: per 512-value unit a count,
// an exclusive scan, then compaction in index order (value i of a unit is
// u*512 + 64m + t, so ballots in m order give ascending ranks).
// ------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_sel_count(const uint64_t *x, const uint64_t *d_n, ms_pred p, uint64_t *cnt,
                                                   uint64_t units_bound)
{
    const uint64_t n = *d_n;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    for (uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index(); u < units_bound; u += nwaves) {
        uint32_t c = 0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const uint64_t i = u * 512 + 64 * m + t;
            const bool take = i < n && pred_eval(p, x[i]);
            c += __popcll(__ballot(take));
        }
        if (t == 0) cnt[u] = c;
    }
}

__global__ __launch_bounds__(WG) void k_sel_compact(const uint64_t *x, const uint64_t *d_n, ms_pred p,
                                                     const uint64_t *off, uint64_t *pos, uint64_t units_bound,
                                                     uint64_t *d_cnt)
{
    const uint64_t n = *d_n;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
    const uint64_t nunits = (n + 511) / 512;
    for (uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index(); u < nunits; u += nwaves) {
        uint64_t o = off[u];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const uint64_t i = u * 512 + 64 * m + t;
            const bool take = i < n && pred_eval(p, x[i]);
            const uint64_t bal = __ballot(take);
            if (take) pos[o + __popcll(bal & ((1ull << t) - 1ull))] = i;
            o += __popcll(bal);
        }
        if (u == nunits - 1 && t == 0) *d_cnt = o;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && n == 0) *d_cnt = 0;
    (void)units_bound;
}

// ------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------
static bool fast_blocks(const ms_colv &v)
{
    // 8-lane indexed layouts whose blocks hold a power-of-two multiple of
    // 512 values (and, for cascades, a power-of-two number of blocks per
    // outer block)
    return (v.g.kind == MS_DYNAMIC_VBP || v.g.kind == MS_FOR || v.g.kind == MS_DELTA) &&
           v.g.leaf == MS_DYNAMIC_VBP && v.g.lanes == 8 && v.g.lg_spb != 0xff &&
           (v.g.kind == MS_DYNAMIC_VBP || v.g.lg_ipo != 0xff);
}

static bool fast_static(const ms_colv &v)
{
    return v.g.kind == MS_STATIC_VBP && v.g.lg_lanes != 0xff && v.g.lanes <= 64;
}

typedef void (*static_kernel_t)(ms_colv, ms_pred, uint64_t *);
typedef void (*static_dec_t)(ms_colv, uint64_t *);

static static_kernel_t static_sum_kernel(uint32_t bw)
{
    switch (bw) {
#define MS_CASE(B) \
    case B: return k_select_sum_static<B>;
        MS_BW_CASES(MS_CASE)
#undef MS_CASE
    default: return nullptr;
    }
}

static static_dec_t static_dec_kernel(uint32_t bw)
{
    switch (bw) {
#define MS_CASE(B) \
    case B: return k_decomp_static<B>;
        MS_BW_CASES(MS_CASE)
#undef MS_CASE
    default: return nullptr;
    }
}

// Sums the per-workgroup partials and the uncompressed rest (select + sum
// over the rest part, the vectorised and scalar loops of
// agg_sum_compr.h:110-146 folded into one) and writes the result.
__global__ __launch_bounds__(WG) void k_sum_finish(ms_colv c, ms_pred p, const uint64_t *part, uint32_t nparts,
                                                   int with_rest, uint64_t *result)
{
    uint64_t acc = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += WG) acc += part[i];
    if (with_rest) {
        const ms_parts pt = parts_of(c);
        const uint64_t *x = (const uint64_t *)(c.data + pt.rest_off);
        for (uint64_t i = threadIdx.x; i < pt.n_rest; i += WG) acc += pred_eval(p, x[i]) ? x[i] : 0;
    }
    __shared__ uint64_t s_part[WAVES];
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t sum = 0;
        for (int w = 0; w < WAVES; w++) sum += s_part[w];
        *result = sum;
    }
}

static bool delta_fast(const ms_colv &v) { return fast_blocks(v) && v.g.dstep == 8; }
static bool delta_shiftable(const ms_colv &v) { return v.g.lg_dstep != 0xff && v.g.dstep <= 8; }

// Deltas of a delta_f column whose inner layout has no fast decoder,
// decoded per value into a temporary.
static int delta_src(const ms_col *in, const ms_colv &v, temps &sc, const uint64_t **arr,
                     const ss_buffers *bufs = nullptr)
{
    *arr = nullptr;
    if (delta_fast(v)) return MS_OK;
    uint64_t *tmp;
    if (bufs) {
        if (in->count_bound > bufs->tmp_values)
            return ms_fail(MS_ERR_CAPACITY, "delta_f column of %llu values without a unit decoder exceeds the "
                           "%llu-value temporary of the caller's buffers",
                           (unsigned long long)in->count_bound, (unsigned long long)bufs->tmp_values);
        tmp = bufs->tmp;
    } else {
        tmp = (uint64_t *)sc.get(in->count_bound * 8 + 64);
    }
    if (!tmp) return ms_fail(MS_ERR_NOMEM, "delta scratch");
    hipLaunchKernelGGL(k_decomp_generic, dim3(grid_for(in->count_bound, WG * 8ull)), dim3(WG), 0, sc.s, v, tmp);
    MS_HIP(hipGetLastError());
    *arr = tmp;
    return MS_OK;
}

int launch_delta_carries(const ms_col *in, temps &sc)
{
    ms_colv v;
    MS_TRY(colv_of(in, v));
    const uint64_t ntb = in->count_bound / DELTA_TILE + 1;
    if (!in->refs || in->refs_cap < ntb * 8) return ms_fail(MS_ERR_CAPACITY, "delta carries capacity");
    const uint64_t *arr;
    MS_TRY(delta_src(in, v, sc, &arr));
    uint64_t *sums = (uint64_t *)sc.get(ntb * 8 * 8 + 64);
    uint64_t *scan = (uint64_t *)sc.get(ntb * 8 * 8 + 64);
    if (!sums || !scan) return ms_fail(MS_ERR_NOMEM, "delta carries scratch");
    const uint32_t S = in->g.dstep;
    const dim3 grid(grid_for(ntb, 1, 65536));
    if (arr) hipLaunchKernelGGL(k_delta_tile_sums<1>, grid, dim3(WG), 0, sc.s, v, arr, sums, ntb);
    else hipLaunchKernelGGL(k_delta_tile_sums<0>, grid, dim3(WG), 0, sc.s, v, arr, sums, ntb);
    MS_HIP(hipGetLastError());
    for (uint32_t j = 0; j < S; j++) MS_TRY(launch_exclusive_scan(sums + j * ntb, scan + j * ntb, ntb, sc));
    hipLaunchKernelGGL(k_delta_carries, dim3(grid_for(ntb * 8, WG)), dim3(WG), 0, sc.s, scan, ntb, S, in->refs);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

static uint32_t units_grid(uint64_t count_bound) { return grid_for(count_bound / 512 + 1, WAVES * 8ull, 2048); }

// Workgroups of `kernel` that the device holds at once (occupancy per CU
// times CUs), queried once per kernel: grid-stride kernels that prefetch
// their next piece of work run best as one resident wave of workgroups.
static uint32_t resident_grid(const void *kernel)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, uint32_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 2048;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(dev, kernel);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0, blocks = 0;
    uint32_t g = 2048;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, WG, 0) == hipSuccess && cus > 0 && blocks > 0)
        g = (uint32_t)(cus * blocks);
    cache[key] = g;
    return g;
}

// Grid of a grid-stride kernel over `items` pieces of one wave each.
static uint32_t stride_grid(const void *kernel, uint64_t items)
{
    const uint32_t r = resident_grid(kernel);
    return grid_for(items, WAVES, r < 8192 ? r : 8192);
}

int launch_select_sum(const ms_col *in, const ms_pred &p, uint64_t *d_result, bool accumulate, hipStream_t s,
                      const ss_buffers *bufs)
{
    ms_colv v;
    MS_TRY(colv_of(in, v));
    if (v.g.kind == MS_BLOCKWISE_INDIVIDUAL) {
        if (bufs) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: select_sum needs temporaries of its own");
        return launch_bwi_select_sum(in, p, d_result, accumulate, s);
    }
    temps sc(s);
    const uint64_t nb = in->count_bound;
    uint32_t grid = 1;
    int with_rest = 1;
    uint64_t *part = bufs ? bufs->part : (uint64_t *)sc.get(SS_MAX_PARTS * 8);
    if (!part) return ms_fail(MS_ERR_NOMEM, "select_sum scratch");
    switch (v.g.kind) {
    case MS_UNCOMPR:
        grid = grid_for(nb, WG * 32ull, 2048);
        hipLaunchKernelGGL(k_select_sum_uncompr, dim3(grid), dim3(WG), 0, s, v, p, part);
        with_rest = 0;
        break;
    case MS_RLE:
        grid = stride_grid((const void *)k_select_sum_rle, in->size_alloc / 16 / RLE_CHUNK + 1);
        hipLaunchKernelGGL(k_select_sum_rle, dim3(grid), dim3(WG), 0, s, v, p, part);
        with_rest = 0;
        break;
    case MS_STATIC_VBP:
        if (fast_static(v)) {
            grid = units_grid(nb);
            hipLaunchKernelGGL(static_sum_kernel(v.g.bw), dim3(grid), dim3(WG), 0, s, v, p, part);
        } else {
            grid = grid_for(nb, WG * 16ull, 2048);
            hipLaunchKernelGGL(k_select_sum_generic, dim3(grid), dim3(WG), 0, s, v, p, part);
        }
        break;
    case MS_DYNAMIC_VBP:
    case MS_FOR:
        if (fast_blocks(v)) {
            const ms_bwtab tab = make_bwtab(p);
            const uint64_t quads = nb / 2048 + 1;   // groups of 4 units
            auto k = v.g.kind == MS_FOR && pred_le_form(p) ? k_select_sum_blocks<true, true>
                   : v.g.kind == MS_FOR                    ? k_select_sum_blocks<true, false>
                                                           : k_select_sum_blocks<false, false>;
            grid = stride_grid((const void *)k, quads);
            hipLaunchKernelGGL(k, dim3(grid), dim3(WG), 0, s, v, p, tab, part);
        } else {
            grid = grid_for(nb, WG * 16ull, 2048);
            hipLaunchKernelGGL(k_select_sum_generic, dim3(grid), dim3(WG), 0, s, v, p, part);
        }
        break;
    case MS_DELTA: {
        const uint64_t *arr;
        MS_TRY(delta_src(in, v, sc, &arr, bufs));
        if (!arr && v.g.lg_spb != 0xff) {
            grid = stride_grid((const void *)k_delta_select_sum, nb / DELTA_TILE + 1);
            hipLaunchKernelGGL(k_delta_select_sum, dim3(grid), dim3(WG), 0, s, v, p, part);
            break;
        }
        grid = grid_for(nb / DELTA_TILE + 1, 2, 2048);
        if (arr) hipLaunchKernelGGL((k_delta_tiles<1, 1>), dim3(grid), dim3(WG), 0, s, v, arr, nullptr, p, part);
        else hipLaunchKernelGGL((k_delta_tiles<1, 0>), dim3(grid), dim3(WG), 0, s, v, arr, nullptr, p, part);
        break;
    }
    case MS_STATIC_VBP_PAD:
    case MS_K_WISE_NS:
    case MS_GROUP_SIMPLE: {
        // these layouts are read through decompression into a temporary
        uint64_t *tmp;
        if (bufs) {
            if (nb > bufs->tmp_values)
                return ms_fail(MS_ERR_CAPACITY, "a column of %llu values exceeds the %llu-value temporary of the "
                               "caller's buffers", (unsigned long long)nb, (unsigned long long)bufs->tmp_values);
            tmp = bufs->tmp;
        } else {
            tmp = (uint64_t *)sc.get(nb * 8 + 64);
        }
        if (!tmp) return ms_fail(MS_ERR_NOMEM, "select_sum temporary");
        MS_TRY(launch_decompress(in, tmp, s));
        grid = grid_for(nb, WG * 32ull, 2048);
        hipLaunchKernelGGL(k_select_sum_values, dim3(grid), dim3(WG), 0, s, (const uint64_t *)tmp, &in->meta->count,
                           p, part);
        with_rest = 0;
        break;
    }
    default:
        return ms_fail(MS_ERR_INVALID, "select_sum: unsupported format");
    }
    MS_HIP(hipGetLastError());
    static_assert(SS_MAX_PARTS >= 8192, "every grid above is capped at 8192 workgroups");
    hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(WG), 0, s, v, p, (const uint64_t *)part, grid, with_rest, d_result);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// Decompressed values are written once, far more of them than any cache
// holds: the unit and delta decoders store them non-temporally (measured:
// dynamic_vbp 334 -> 247 us, delta 400 -> 232 us for 2^27 values with the
// paired 16-byte stores).  rle expansion measured slower that way
// (339 -> 421 us) and stores normally.
#ifndef MS_DECOMP_NT
#define MS_DECOMP_NT 1
#endif
static constexpr bool DECOMP_NT = MS_DECOMP_NT != 0;

int launch_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s, ms_dev_meta *om)
{
    ms_colv v;
    MS_TRY(colv_of(in, v));
    const uint64_t nb = in->count_bound;
    if (nb == 0) {
        if (om) hipLaunchKernelGGL(k_set_meta_uncompr, dim3(1), dim3(1), 0, s, (const ms_dev_meta *)in->meta, om);
        return om ? (hipGetLastError() == hipSuccess ? MS_OK : ms_fail(MS_ERR_HIP, "decompress meta")) : MS_OK;
    }
    const uint32_t grid_units = grid_for(nb / 512 + 1, WAVES * 4ull, 8192);
    const uint64_t quads = nb / 2048 + 1;   // groups of 4 units
    switch (v.g.kind) {
    case MS_UNCOMPR:
        break;
    case MS_RLE: {
        temps sc(s);
        const uint64_t nch_bound = (in->size_alloc / 16 + 1) / RLE_XCH + 1;
        uint64_t *clen = (uint64_t *)sc.get(nch_bound * 8);
        uint64_t *cst = (uint64_t *)sc.get(nch_bound * 8);
        uint64_t *d_nch = (uint64_t *)sc.get(64);
        if (!clen || !cst || !d_nch) return ms_fail(MS_ERR_NOMEM, "rle scratch");
        hipLaunchKernelGGL(k_rle_chunk_lens, dim3(grid_for(nch_bound, 1, 8192)), dim3(WG), 0, s, v, clen, d_nch);
        MS_TRY(launch_exclusive_scan_n(clen, cst, d_nch, 0, nch_bound, sc));
        auto k = k_rle_expand_chunks<false>;
        hipLaunchKernelGGL(k, dim3(stride_grid((const void *)k, nch_bound)), dim3(WG), 0, s, v, cst, d_out);
        if (om) hipLaunchKernelGGL(k_set_meta_uncompr, dim3(1), dim3(1), 0, s, (const ms_dev_meta *)in->meta, om);
        MS_HIP(hipGetLastError());
        return MS_OK;
    }
    case MS_STATIC_VBP:
        if (fast_static(v) && v.g.lanes == 8) {
            auto k = k_decomp_units<DL_STATIC8, DECOMP_NT>;
            hipLaunchKernelGGL(k, dim3(stride_grid((const void *)k, quads)), dim3(WG), 0, s, v, d_out, om);
            MS_HIP(hipGetLastError());
            return MS_OK;   // rest and meta written by the kernel
        } else if (fast_static(v)) {
            hipLaunchKernelGGL(static_dec_kernel(v.g.bw), dim3(grid_units), dim3(WG), 0, s, v, d_out);
        } else {
            hipLaunchKernelGGL(k_decomp_generic, dim3(grid_for(nb, WG * 8ull)), dim3(WG), 0, s, v, d_out);
        }
        break;
    case MS_DYNAMIC_VBP:
    case MS_FOR:
        if (fast_blocks(v)) {
            auto k = v.g.kind == MS_FOR ? k_decomp_units<DL_FOR, DECOMP_NT> : k_decomp_units<DL_BLOCKS, DECOMP_NT>;
            hipLaunchKernelGGL(k, dim3(stride_grid((const void *)k, quads)), dim3(WG), 0, s, v, d_out, om);
            MS_HIP(hipGetLastError());
            return MS_OK;   // rest and meta written by the kernel
        } else {
            hipLaunchKernelGGL(k_decomp_generic, dim3(grid_for(nb, WG * 8ull)), dim3(WG), 0, s, v, d_out);
        }
        break;
    case MS_DELTA: {
        temps sc(s);
        const uint64_t *arr;
        MS_TRY(delta_src(in, v, sc, &arr));
        if (!arr && v.g.lg_spb != 0xff) {
            auto k = k_delta_decomp<DECOMP_NT>;
            hipLaunchKernelGGL(k, dim3(stride_grid((const void *)k, nb / DELTA_TILE + 1)), dim3(WG), 0, s, v, d_out,
                               om);
            MS_HIP(hipGetLastError());
            return MS_OK;   // rest and meta written by the kernel
        }
        const dim3 grid(grid_for(nb / DELTA_TILE + 1, 1, 65536));
        if (arr) hipLaunchKernelGGL((k_delta_tiles<0, 1>), grid, dim3(WG), 0, s, v, arr, d_out, ms_pred{}, nullptr);
        else hipLaunchKernelGGL((k_delta_tiles<0, 0>), grid, dim3(WG), 0, s, v, arr, d_out, ms_pred{}, nullptr);
        break;
    }
    case MS_STATIC_VBP_PAD:
    case MS_K_WISE_NS:
    case MS_GROUP_SIMPLE:
        MS_TRY(launch_xfmt_decompress(in, d_out, s));
        break;
    case MS_BLOCKWISE_INDIVIDUAL:
        MS_TRY(launch_bwi_decompress(in, d_out, s));   // copies its rest itself
        if (om) hipLaunchKernelGGL(k_set_meta_uncompr, dim3(1), dim3(1), 0, s, (const ms_dev_meta *)in->meta, om);
        MS_HIP(hipGetLastError());
        return MS_OK;
    default:
        return ms_fail(MS_ERR_INVALID, "decompress: unsupported format");
    }
    MS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_copy_rest, dim3(grid_for(nb, WG * 4ull, 4096)), dim3(WG), 0, s, v, d_out);
    if (om) hipLaunchKernelGGL(k_set_meta_uncompr, dim3(1), dim3(1), 0, s, (const ms_dev_meta *)in->meta, om);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

__global__ void k_count_of(const ms_dev_meta *m, uint64_t *dst) { *dst = m->count; }

__global__ void k_units_of(const ms_dev_meta *m, uint64_t *dst) { *dst = (m->count + 511) / 512; }

// Flags of unit u's 512 values (bit t of the result's word m = value
// 64 * (t >> 3) + 8 * m + (t & 7) of the unit, lane-major per thread):
// decoded from the compressed part, or read raw from the rest.
// Index entry and for_f reference of unit u (left in vector registers, so
// that a caller can request them one unit ahead and wait only at use).
__device__ __forceinline__ void unit_entry(const ms_colv &c, uint64_t ncu, uint64_t u, uint64_t &e, uint64_t &r)
{
    e = r = 0;
    if (u < ncu) {
        const uint64_t b = u >> c.g.lg_spb;
        e = gload(c.index + b);
        if (c.g.kind == MS_FOR) r = gload(c.refs + (b >> c.g.lg_ipo));
    }
}

__device__ __forceinline__ void unit_flags(const ms_colv &c, const ms_parts &pt, const ms_pred &p, uint64_t u,
                                           uint64_t e_v, uint64_t r_v, uint32_t t, uint64_t (&B)[8])
{
    const uint32_t j = t & 7, gl = t >> 3;
    const uint64_t ncu = pt.n_compr >> 9;
    uint64_t d[8];
    bool ok[8];
    if (u < ncu) {
        const uint64_t sub = u & ((1ull << c.g.lg_spb) - 1);
        const uint64_t e = uniform64(e_v);
        const uint64_t ref = c.g.kind == MS_FOR ? uniform64(r_v) : 0ull;
        dec_dispatch((uint32_t)(e & 0xff), (const uint64_t *)(c.data + (e >> 8)), 8, j, sub * 8 + gl, d);
#pragma unroll
        for (int m = 0; m < 8; m++) {
            d[m] += ref;
            ok[m] = true;
        }
    } else {
        // the raw rest: value i of the unit is rest value (u - ncu) * 512 + i
        const uint64_t *x = (const uint64_t *)(c.data + pt.rest_off);
        const uint64_t r0 = (u - ncu) * 512;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const uint64_t r = r0 + 64 * gl + 8 * m + j;
            ok[m] = r < pt.n_rest;
            d[m] = ok[m] ? gload(x + r) : 0ull;
        }
    }
#pragma unroll
    for (int m = 0; m < 8; m++) B[m] = __ballot(ok[m] && pred_eval(p, d[m]));
}


// Decodes and tests every unit once: its 8 ballots go to flags[8 u ..]
// (64 bytes per 512 values) and its selected count to cnt[u].
__global__ __launch_bounds__(WG) void k_sel_count_blocks(ms_colv c, ms_pred p, uint64_t *d_units, uint64_t *cnt,
                                                         uint64_t *flags)
{
    const ms_parts pt = parts_of(c);
    const uint64_t nunits = (pt.n + 511) / 512;
    const uint64_t ncu = pt.n_compr >> 9;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *d_units = nunits;   // for the scan and the compaction after this kernel
        cnt[nunits] = 0;     // the scan's total slot
    }
    uint64_t e_n, r_n;   // the next unit's entry, requested one unit ahead
    unit_entry(c, ncu, u, e_n, r_n);
    for (; u < nunits; u += W) {
        const uint64_t e_v = e_n, r_v = r_n;
        unit_entry(c, ncu, u + W, e_n, r_n);
        uint64_t B[8];
        unit_flags(c, pt, p, u, e_v, r_v, t, B);
        uint32_t n = 0;
        uint64_t mine = 0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
            n += __popcll(B[m]);
            if (t == (uint32_t)m) mine = B[m];
        }
        if (t < 8) flags[8 * u + t] = mine;
        if (t == 0) cnt[u] = n;
    }
}

// Writes the selected positions of every unit from its stored ballots
// (no second decode).
__global__ __launch_bounds__(WG) void k_sel_compact_blocks(const uint64_t *flags, const uint64_t *d_units,
                                                           const uint64_t *off, uint64_t *pos, uint64_t *d_cnt)
{
    const uint64_t nunits = *d_units;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t W = (uint64_t)gridDim.x * WAVES;
    const uint64_t below = t ? (~0ull >> (64 - t)) : 0ull;
    uint64_t u = (uint64_t)blockIdx.x * WAVES + wave_index();
    // One load per lane requests a unit: lanes 0..7 its ballots, lane 8 its
    // output offset (the other lanes re-read ballot 0).  The next unit is
    // requested one unit ahead; past the end the current unit again.
    const auto request = [&](uint64_t uu) {
        return gload(t < 8 ? flags + 8 * uu + t : (t == 8 ? off + uu : flags + 8 * uu));
    };
    uint64_t my_n = u < nunits ? request(u) : 0ull;
    for (; u < nunits; u += W) {
        const uint64_t my = my_n;
        my_n = request(u + W < nunits ? u + W : u);
        // value 64 r + t of the unit is bit 8 r + (t & 7) of ballot t / 8:
        // lane t takes that ballot, and row r's flags are one ballot of its
        // bit r (a transpose of the 8 x 8 byte matrix, one step per row)
        const uint64_t mine = __shfl(my, (int)(t >> 3), MS_WAVE);
        uint64_t o = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(my >> 32), 8) << 32) |
                     __builtin_amdgcn_readlane((uint32_t)my, 8);
#pragma unroll
        for (uint32_t r = 0; r < 8; r++) {
            const bool take = (mine >> (8 * r + (t & 7))) & 1;
            const uint64_t F = __ballot(take);
            if (take) pos[o + __popcll(F & below)] = u * 512 + 64 * r + t;
            o += __popcll(F);
        }
        if (u == nunits - 1 && t == 0) *d_cnt = o;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && nunits == 0) *d_cnt = 0;
}

bool select_flags_ok(const ms_col *in)
{
    ms_colv v;
    if (colv_of(in, v) != MS_OK) return false;
    return (v.g.kind == MS_DYNAMIC_VBP || v.g.kind == MS_FOR) && fast_blocks(v);
}

// The count pass of the block-layout select and the scan of its unit
// counts (with the total at uoff[units]); no compaction.
int launch_select_flags(const ms_col *in, const ms_pred &p, sel_flags &sf, temps &sc)
{
    hipStream_t s = sc.s;
    const uint64_t nb = in->count_bound;
    ms_colv v;
    MS_TRY(colv_of(in, v));
    if (!((v.g.kind == MS_DYNAMIC_VBP || v.g.kind == MS_FOR) && fast_blocks(v)))
        return ms_fail(MS_ERR_INVALID, "select flags: layout without a block decoder");
    const uint64_t units = nb / 512 + 1;
    uint64_t *d_u = (uint64_t *)sc.get(64);
    uint64_t *cnt = (uint64_t *)sc.get(units * 8 + 64);
    uint64_t *off = (uint64_t *)sc.get(units * 8 + 64);
    uint64_t *flags = (uint64_t *)sc.get(units * 64 + 64);
    if (!d_u || !cnt || !off || !flags) return ms_fail(MS_ERR_NOMEM, "select scratch");
    hipLaunchKernelGGL(k_sel_count_blocks, dim3(units_grid(nb)), dim3(WG), 0, s, v, p, d_u, cnt, flags);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan_n(cnt, off, d_u, 1, units + 1, sc));
    sf.flags = flags;
    sf.uoff = off;
    sf.d_units = d_u;
    sf.units_bound = units;
    return MS_OK;
}

int launch_select_positions(const ms_col *in, const ms_pred &p, uint64_t *d_pos, uint64_t *d_cnt, temps &sc)
{
    hipStream_t s = sc.s;
    const uint64_t nb = in->count_bound;
    ms_colv v;
    MS_TRY(colv_of(in, v));
    if ((v.g.kind == MS_DYNAMIC_VBP || v.g.kind == MS_FOR) && fast_blocks(v)) {
        // decode, test and compact straight from the compressed units
        const uint64_t units = nb / 512 + 1;
        uint64_t *d_u = (uint64_t *)sc.get(64);
        uint64_t *cnt = (uint64_t *)sc.get(units * 8 + 64);
        uint64_t *off = (uint64_t *)sc.get(units * 8 + 64);
        uint64_t *flags = (uint64_t *)sc.get(units * 64 + 64);
        if (!d_u || !cnt || !off || !flags) return ms_fail(MS_ERR_NOMEM, "select scratch");
        const dim3 grid(units_grid(nb));
        hipLaunchKernelGGL(k_sel_count_blocks, grid, dim3(WG), 0, s, v, p, d_u, cnt, flags);
        MS_HIP(hipGetLastError());
        MS_TRY(launch_exclusive_scan_n(cnt, off, d_u, 0, units, sc));
        hipLaunchKernelGGL(k_sel_compact_blocks, grid, dim3(WG), 0, s, (const uint64_t *)flags,
                           (const uint64_t *)d_u, (const uint64_t *)off, d_pos, d_cnt);
        MS_HIP(hipGetLastError());
        return MS_OK;
    }
    const uint64_t *x;
    if (in->g.kind == MS_UNCOMPR) {
        x = (const uint64_t *)in->data;
    } else {
        uint64_t *tmp = (uint64_t *)sc.get(nb * 8 + 64);
        if (!tmp) return ms_fail(MS_ERR_NOMEM, "select scratch");
        MS_TRY(launch_decompress(in, tmp, s));
        x = tmp;
    }
    uint64_t *d_n = (uint64_t *)sc.get(64);
    const uint64_t units = nb / 512 + 1;
    uint64_t *cnt = (uint64_t *)sc.get(units * 8 + 64);
    uint64_t *off = (uint64_t *)sc.get(units * 8 + 64);
    if (!d_n || !cnt || !off) return ms_fail(MS_ERR_NOMEM, "select scratch");
    uint64_t *d_u = (uint64_t *)sc.get(64);
    if (!d_u) return ms_fail(MS_ERR_NOMEM, "select scratch");
    hipLaunchKernelGGL(k_count_of, dim3(1), dim3(1), 0, s, in->meta, d_n);
    hipLaunchKernelGGL(k_units_of, dim3(1), dim3(1), 0, s, in->meta, d_u);
    hipLaunchKernelGGL(k_sel_count, dim3(grid_for(units, WAVES * 2ull)), dim3(WG), 0, s, x, d_n, p, cnt, units);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan_n(cnt, off, d_u, 0, units, sc));
    hipLaunchKernelGGL(k_sel_compact, dim3(grid_for(units, WAVES * 2ull)), dim3(WG), 0, s, x, d_n, p, off, d_pos,
                       units, d_cnt);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------
// Exclusive scan of the first min(*d_n + extra, bound) values, sized from
// the device count: G workgroups each take one contiguous chunk (sum,
// then the chunk scan on the sum of the sums before it).  Arrays allocated for
// a host bound are scanned only as far as they are used; without a device
// count (d_n null) all `bound` values are scanned.  Counts are 64-bit
// throughout.
// ------------------------------------------------------------------
__device__ __forceinline__ void chunk_of(const uint64_t *d_n, uint64_t extra, uint64_t bound, uint64_t &lo,
                                         uint64_t &hi)
{
    uint64_t n = d_n ? *d_n + extra : bound;
    if (n > bound) n = bound;
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    lo = (uint64_t)blockIdx.x * chunk;
    hi = lo + chunk < n ? lo + chunk : n;
    if (lo > hi) lo = hi;
}

__device__ __forceinline__ uint64_t block_sum(uint64_t v)
{
    __shared__ uint64_t s_w[WAVES];
    v = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    uint64_t t = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) t += s_w[w];
    __syncthreads();
    return t;
}

// Exclusive scan over the workgroup; *total receives the sum.
__device__ __forceinline__ uint64_t block_exscan(uint64_t v, uint64_t *total)
{
    __shared__ uint64_t s_w[WAVES];
    const uint32_t t = threadIdx.x & 63;
    uint64_t incl = v;
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(incl, o, MS_WAVE);
        if (t >= o) incl += y;
    }
    if (t == 63) s_w[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < WAVES; w++) {
        const uint64_t x = s_w[w];
        if (w < (threadIdx.x >> 6)) before += x;
        all += x;
    }
    __syncthreads();
    *total = all;
    return before + incl - v;
}

__global__ __launch_bounds__(WG) void k_scan_part(const uint64_t *in, const uint64_t *d_n, uint64_t extra,
                                                  uint64_t bound, uint64_t *part)
{
    uint64_t lo, hi;
    chunk_of(d_n, extra, bound, lo, hi);
    uint64_t acc = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += WG) acc += in[i];
    acc = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(WG) void k_scan_apply(const uint64_t *in, uint64_t *out, const uint64_t *d_n,
                                                   uint64_t extra, uint64_t bound, const uint64_t *part)
{
    uint64_t lo, hi;
    chunk_of(d_n, extra, bound, lo, hi);
    // the chunk's offset: the sums of the chunks before it (at most 1024)
    uint64_t before = 0;
    for (uint32_t k = threadIdx.x; k < blockIdx.x; k += WG) before += part[k];
    uint64_t carry = block_sum(before);
    for (uint64_t base = lo; base < hi; base += WG) {
        const uint64_t i = base + threadIdx.x;
        const uint64_t v = i < hi ? in[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_exscan(v, &tot);
        if (i < hi) out[i] = carry + ex;
        carry += tot;
    }
}

int launch_exclusive_scan_n(const uint64_t *d_in, uint64_t *d_out, const uint64_t *d_n, uint64_t extra,
                            uint64_t bound, temps &sc)
{
    if (bound == 0) return MS_OK;
    const uint32_t G = grid_for(bound, 4096, 1024);
    uint64_t *part = (uint64_t *)sc.get(G * 8 + 64);
    if (!part) return ms_fail(MS_ERR_NOMEM, "scan scratch");
    hipLaunchKernelGGL(k_scan_part, dim3(G), dim3(WG), 0, sc.s, d_in, d_n, extra, bound, part);
    hipLaunchKernelGGL(k_scan_apply, dim3(G), dim3(WG), 0, sc.s, d_in, d_out, d_n, extra, bound,
                       (const uint64_t *)part);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

int launch_exclusive_scan(const uint64_t *d_in, uint64_t *d_out, uint64_t n, temps &sc)
{
    return launch_exclusive_scan_n(d_in, d_out, nullptr, 0, n, sc);
}

}  // namespace ms
