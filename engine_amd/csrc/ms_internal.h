// ms_internal.h -- structures shared by the host layer and the kernels.
//
// A device column keeps the reference's byte image (column.h:76-109) plus:
//   * ms_dev_meta  : count / sizeUsed / sizeCompr and a status word, in
//                    device memory, so producers can set them without a
//                    host round trip;
//   * a block index: one 64-bit entry per innermost vertical-bit-packed
//                    block of the compressed part, (payload byte offset << 8)
//                    | bit width.  It is derived data, not part of the image;
//   * for for_f     : a copy of each FOR block's reference value;
//   * for delta_f   : per MS_DELTA_TILE values, the lane sums at the tile's
//                     start (8 slots per tile).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/morphstore_hip.h"

#define MS_HD __host__ __device__ __forceinline__

// delta_f columns keep, per tile of this many values of the compressed
// part, the running sums of the delta lanes at the tile's start.
#define MS_DELTA_TILE 2048ull

struct ms_dev_meta {
    uint64_t count;
    uint64_t size_used;
    uint64_t size_compr;
    uint64_t status;      // MS_OK or the first error a producing kernel saw
    uint64_t aux[4];      // scratch slots for multi-kernel producers
};

// Innermost packing geometry of a format: what the kernels decode.
struct ms_geom {
    uint32_t kind;        // outer kind
    uint32_t leaf;        // kind of the innermost layout (uncompr/static/dvbp)
    uint32_t bw;          // static_vbp bit width (leaf == MS_STATIC_VBP)
    uint32_t lanes;       // lanes of the innermost layout
    uint64_t bs;          // outer m_BlockSize (1 for uncompr, 0 for rle)
    uint64_t inner_bs;    // innermost block size in values (vbp block)
    uint64_t leaf_page;   // dynamic_vbp blocks per page (innermost)
    uint64_t outer_block; // delta/for block size
    uint64_t outer_page;  // for: blocks per page
    uint64_t gap;         // 2 * F::get_size_max_byte(F::m_BlockSize)
    uint32_t dstep;       // delta_f: step of the differences
    // log2 of ratios the kernels divide by (0xff: not a power of two)
    uint8_t lg_spb;       // 512-value units per innermost block
    uint8_t lg_ipo;       // innermost blocks per outer (delta/for) block
    uint8_t lg_lanes;     // innermost lanes
    uint8_t lg_dstep;     // delta step
};

MS_HD uint8_t ms_lg2(uint64_t x)
{
    if (x == 0 || (x & (x - 1))) return 0xff;
    uint8_t l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

MS_HD uint64_t ms_align64(uint64_t x) { return (x + 63ull) & ~63ull; }

// F::get_size_max_byte(n) for n a multiple of F::m_BlockSize
// (static_vbp.h:63-65, dynamic_vbp.h:106-118, delta.h:92-102, for.h:100-112).
MS_HD uint64_t ms_leaf_size_max(uint32_t kind, uint32_t bw, uint64_t block, uint64_t page, uint64_t n)
{
    if (kind == MS_UNCOMPR) return 8 * n;
    if (kind == MS_STATIC_VBP) return n * bw / 8;
    return 8 * n + (n / block) * page;
}

// get_size_max_byte of the formats below (vbp_padding.h:60-63,
// k_wise_ns.h:126-130, group_simple.h:93-121) for n a multiple of the
// block size.
MS_HD uint64_t ms_pad_size_max(uint32_t bw, uint32_t step, uint64_t n)
{
    const uint64_t bs = (uint64_t)(64 / bw) * step;
    return n / bs * step * 8 * 10;
}
MS_HD uint64_t ms_kwns_size_max(uint64_t block, uint64_t n) { return 8 * n + n / block; }
MS_HD uint64_t ms_gs_size_max(uint32_t E, uint64_t n)
{
    const uint64_t calls = (n + 1023) / 1024;   // the reference's fixed 1024-value call bound
    const uint64_t sels = n / E / 2;
    return calls * (8 + 1) + (sels + 1) / 2 * 2 + calls * (8ull * E - 1) + 8 * n;
}

// Values of a k_wise_ns_f column per block-index entry (32 pairs).
#define MS_KWNS_PER_ENTRY 64ull

// blockwise_individual_helper::get_size_max_byte_single_block
// (blockwise_individual.h:57-79) over the default formats that
// default_formats.h defines (uncompr_f, static_vbp_f<vbp_l<64, 8>>,
// dynamic_vbp_f<512, 64, 8>, group_simple_f<8, uint64_t, 64>, delta_f and
// for_f over that dynamic_vbp_f), for B a multiple of 1024.
MS_HD uint64_t ms_bwi_single_max(uint64_t B)
{
    uint64_t m = 8 * B;
    const uint64_t dvbp = ms_leaf_size_max(MS_DYNAMIC_VBP, 0, 512, 64, B);
    const uint64_t gs = ms_gs_size_max(8, B);
    const uint64_t forf = (B + 8 * 1024 - 1) / (8 * 1024) * 64 + B / 1024 * ms_leaf_size_max(MS_DYNAMIC_VBP, 0, 512, 64, 1024);
    m = dvbp > m ? dvbp : m;
    m = gs > m ? gs : m;
    m = forf > m ? forf : m;
    return m;
}

// blockwise_individual_f::get_size_max_byte (blockwise_individual.h:141-153):
// per block the largest single block plus the 64-byte vector of code and
// padding.
MS_HD uint64_t ms_bwi_size_max(uint64_t B, uint64_t n) { return n / B * (ms_bwi_single_max(B) + 64); }

MS_HD ms_geom ms_make_geom(const ms_format_desc &f)
{
    ms_geom g = {};
    g.kind = f.kind;
    switch (f.kind) {
    case MS_UNCOMPR:
        g.leaf = MS_UNCOMPR; g.lanes = 1; g.bs = 1; g.inner_bs = 1; g.gap = 0;
        break;
    case MS_STATIC_VBP:
        g.leaf = MS_STATIC_VBP; g.bw = f.bw; g.lanes = f.step; g.bs = 64ull * f.step; g.inner_bs = g.bs;
        g.gap = 2 * ms_leaf_size_max(MS_STATIC_VBP, f.bw, 0, 0, g.bs);
        break;
    case MS_DYNAMIC_VBP:
        g.leaf = MS_DYNAMIC_VBP; g.lanes = f.step; g.bs = f.block; g.inner_bs = f.block; g.leaf_page = f.page;
        g.gap = 2 * ms_leaf_size_max(MS_DYNAMIC_VBP, 0, f.block, f.page, f.block);
        break;
    case MS_DELTA:
    case MS_FOR: {
        // an uncompressed inner part is indexed as one 64-bit-wide block per
        // outer block (vbp_l<64, S> has the uncompressed layout)
        g.leaf = f.inner_kind; g.bw = f.inner_kind == MS_UNCOMPR ? 64 : f.inner_bw;
        g.lanes = f.inner_kind == MS_UNCOMPR ? 1 : f.inner_step;
        g.inner_bs = f.inner_kind == MS_UNCOMPR ? f.block
                   : f.inner_kind == MS_STATIC_VBP ? 64ull * f.inner_step : f.inner_block;
        g.dstep = f.kind == MS_DELTA ? f.step : 0;
        g.leaf_page = f.inner_page;
        g.bs = f.block; g.outer_block = f.block; g.outer_page = f.page;
        const uint64_t inner_max = ms_leaf_size_max(f.inner_kind, f.inner_bw, f.inner_block, f.inner_page, f.block);
        if (f.kind == MS_DELTA) g.gap = 2 * inner_max;
        else {
            const uint64_t pages = (f.block + (uint64_t)f.page * f.block - 1) / ((uint64_t)f.page * f.block);
            g.gap = 2 * (pages * 8ull * f.page + inner_max);
        }
        break;
    }
    case MS_STATIC_VBP_PAD:
        g.leaf = MS_STATIC_VBP_PAD; g.bw = f.bw; g.lanes = f.step; g.bs = (uint64_t)(64 / f.bw) * f.step;
        g.inner_bs = g.bs; g.gap = 2 * ms_pad_size_max(f.bw, f.step, g.bs);
        break;
    case MS_K_WISE_NS:
        g.leaf = MS_K_WISE_NS; g.lanes = f.block; g.bs = f.block; g.inner_bs = MS_KWNS_PER_ENTRY;
        g.gap = 2 * ms_kwns_size_max(f.block, f.block);
        break;
    case MS_GROUP_SIMPLE:
        g.leaf = MS_GROUP_SIMPLE; g.lanes = f.step; g.bs = f.step; g.inner_bs = f.step;
        g.gap = 2 * ms_gs_size_max(f.step, f.step);
        break;
    case MS_BLOCKWISE_INDIVIDUAL:
        g.leaf = MS_BLOCKWISE_INDIVIDUAL; g.lanes = 8; g.bs = f.block; g.inner_bs = f.block;
        g.gap = 2 * ms_bwi_size_max(f.block, f.block);
        break;
    default:   // MS_RLE: no block size, no rest
        g.leaf = MS_RLE; g.lanes = 1; g.bs = 0; g.inner_bs = 0; g.gap = 0;
        break;
    }
    g.lg_spb = (g.inner_bs && g.inner_bs % 512 == 0) ? ms_lg2(g.inner_bs / 512) : 0xff;
    g.lg_ipo = (g.outer_block && g.inner_bs && g.outer_block % g.inner_bs == 0) ? ms_lg2(g.outer_block / g.inner_bs)
                                                                               : 0xff;
    g.lg_lanes = ms_lg2(g.lanes);
    g.lg_dstep = ms_lg2(g.dstep);
    return g;
}

// create_data_uncompr_start (column.h:182-196) relative to the 64-byte
// aligned base of the image.
MS_HD uint64_t ms_rest_offset(const ms_geom &g, uint64_t compr_end)
{
    return ms_align64(compr_end + g.gap);
}

// get_count_values_uncompr (column.h:205-214).
MS_HD uint64_t ms_count_rest(const ms_geom &g, uint64_t size_used, uint64_t size_compr)
{
    if (size_compr == size_used) return 0;
    return (size_used - ms_rest_offset(g, size_compr)) / 8;
}

// Device-side view of a column passed to kernels by value.
struct ms_colv {
    const uint8_t *data;
    const ms_dev_meta *meta;
    const uint64_t *index;   // per innermost block: (offset << 8) | bw
    const uint64_t *refs;    // for_f: reference per FOR block; delta_f: tile carries
    ms_geom g;
};

// Derived numbers every reader needs; computed from the device meta.
struct ms_parts {
    uint64_t n;        // count
    uint64_t n_rest;   // values in the uncompressed rest
    uint64_t n_compr;  // values in the compressed part
    uint64_t rest_off; // byte offset of the rest
};

MS_HD ms_parts ms_parts_of(const ms_geom &g, uint64_t count, uint64_t size_used, uint64_t size_compr)
{
    ms_parts p;
    p.n = count;
    if (g.kind == MS_RLE) { p.n_rest = 0; p.n_compr = count; p.rest_off = size_used; return p; }
    p.n_rest = ms_count_rest(g, size_used, size_compr);
    p.n_compr = count - p.n_rest;
    p.rest_off = g.kind == MS_UNCOMPR ? 0 : ms_rest_offset(g, size_compr);
    if (g.kind == MS_UNCOMPR) { p.n_rest = count; p.n_compr = 0; }
    return p;
}

// blockwise_individual_f columns: where each block's view lies.  Views
// are listed per kind (codes 65, 67, 69, 66, 71, then uncompr_f and
// static_vbp_f); the block index entries, for_f references and delta_f
// carries of codes 65, 67 and 69 are stored kind after kind, in list
// order, so that the views of one of these codes read as one virtual
// column.  `slot` is a view's position in its kind's list.
enum { BWK_DVBP = 0, BWK_DELTA = 1, BWK_FOR = 2, BWK_GS = 3, BWK_RLE = 4, BWK_PLAIN = 5, BWK_N = 6 };
struct ms_bwi_view {
    uint64_t off;     // payload offset in the image (64-byte aligned)
    uint64_t size;    // payload bytes
    uint32_t code;    // ms_format_code
    uint32_t slot;
};
struct ms_bwi {
    uint64_t n, nb, B, C;          // values, blocks, block size, sizeCompr
    std::vector<ms_bwi_view> v;
    uint64_t lcount[BWK_N], loff[BWK_N];
    uint64_t idx_off[3], refs_off[3];
};

// Host-side handle behind ms_col*.
struct ms_col {
    ms_format_desc f;
    ms_geom g;
    uint8_t *data;           // device image, size_alloc bytes (+ slack)
    uint64_t size_alloc;
    ms_dev_meta *meta;       // device
    uint64_t *index;         // device, index_cap entries (or null)
    uint64_t index_cap;
    uint64_t *refs;          // device, refs_cap entries (for_f) or null
    uint64_t refs_cap;
    uint64_t count_bound;    // host-known upper bound of count
    hipStream_t stream;      // stream of creation (frees are ordered on it)
    struct ms_plan *arena;   // non-null: memory belongs to this plan's arena
    uint64_t prepared_bytes; // prepared for random access: compressed rest bytes at sizeCompr
    bool prepared;
    ms_bwi *bwi;             // blockwise_individual_f: the views (host)
    ms_dev_meta *vmeta;      // blockwise_individual_f: meta of every view, then of the 3 groups
};
