// ms_batch.hip -- fused select + agg_sum over many columns of uneven length
// and mixed formats in one persistent launch.
//
// The host cuts every enqueued column into work items (a run of 512-value
// units, of 2048-value delta tiles, of rle runs, or of raw values), in
// entry order.  An item carries what its wave needs (pointers, predicate,
// the block-index entry of its first unit), so starting one costs a single
// scalar load, issued one item ahead.  The persistent kernel's W waves take
// the items round-robin: wave w does items w, w + W, w + 2W, ...; so every
// column is spread over all waves and neighbouring waves stream
// neighbouring bytes.  A wave's items of one entry are consecutive in its
// sequence: it sums them in registers and writes one partial per entry it
// touched, into the slot of its first item of that entry -- for entry e
// with items [lo, hi) exactly the slots [lo, min(hi, lo + W)).  A second
// kernel folds those.  The item sizes per layout and the workgroups per CU
// are picked by ms_batch_tune.
#include <algorithm>
#include <vector>

#include "ms_units.hpp"

namespace ms {

enum { BI_STATIC = 0, BI_BLOCKS = 1, BI_FOR = 2, BI_DELTA = 3, BI_RLE = 4, BI_RAW = 5, BI_GENERIC = 6, BI_DELTA_GEN = 7 };

struct ms_bentry {
    ms_colv c;
    ms_pred p;
    uint64_t item_lo, item_hi;   // the entry's items in the queue
    uint32_t le;                 // p is x <= hi (pred_le_form)
    ms_bwtab tab;                // p specialised per bit width for codes = values
};

// A work item (96 bytes: one scalar load of 6 dwordx4).
struct ms_bitem {
    const uint8_t *data;     // image (BI_RAW: the uncompressed values)
    const uint64_t *index;   // block index (BI_BLOCKS, BI_FOR, BI_DELTA)
    const uint64_t *refs;    // for_f references, delta_f tile carries
    uint64_t first;          // first unit / tile / run / value of the item
    uint64_t ulim;           // BI_DELTA: units of the column's compressed part
    uint64_t tag;            // timed runs: the view (entry, or block of a blockwise_individual_f entry) it belongs to
    uint64_t plo, phi;       // predicate range
    uint32_t entry;
    uint32_t kind_n;         // kind << 28 | number of units / tiles / runs / values
    uint32_t aux;            // BI_STATIC: bit width
    uint8_t lg_spb, lg_ipo;  // log2 units per block, blocks per for_f block
    uint8_t pflags;          // bit 0: negated, bit 1: empty, bit 2: x <= hi form
    uint8_t pad0;
    uint64_t pad1[2];
};
static_assert(sizeof(ms_bitem) == 96, "work item layout");

// The item's n <= 64 units: lane k holds unit u0 + k's index entry (and
// for_f reference), loaded together.
template <bool IS_FOR, int NU>
__device__ __forceinline__ void item_blocks(const ms_bentry &en, const ms_bitem &it, const ms_pred &p, uint32_t n,
                                            uint32_t t, uint32_t j, uint32_t gl, uint64_t &acc)
{
    const uint32_t lg_spb = it.lg_spb, lg_ipo = it.lg_ipo;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    const uint64_t u0 = it.first;
    uint64_t my_e = 0, my_r = 0;
    if (t < n) {
        my_e = gload(it.index + ((u0 + t) >> lg_spb));
        if (IS_FOR) my_r = gload(it.refs + (((u0 + t) >> lg_spb) >> lg_ipo));
    }
    const bool le = IS_FOR && pred_le_form(p);
    for (uint32_t k = 0; k < n; k += NU) {
        const uint32_t cnt = n - k < (uint32_t)NU ? n - k : (uint32_t)NU;
        uint64_t e[4] = {0, 0, 0, 0}, r[4] = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < NU; u++) {
            e[u] = readlane64(my_e, k + u);
            r[u] = IS_FOR ? readlane64(my_r, k + u) : 0ull;
        }
        if (le) vbp_units_ss<IS_FOR, true, NU>(it.data, e, r, cnt, u0 + k, mask, p, en.tab, j, gl, acc);
        else vbp_units_ss<IS_FOR, false, NU>(it.data, e, r, cnt, u0 + k, mask, p, en.tab, j, gl, acc);
    }
}

// The item's n tiles of 4 units (n <= 16), index entries as above.
__device__ __forceinline__ void item_delta(const ms_bitem &it, const ms_pred &p, uint32_t n, uint32_t t, uint32_t j,
                                           uint32_t gl, uint64_t &acc)
{
    const uint32_t lg_spb = it.lg_spb;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    const uint64_t u0 = it.first * 4;
    const uint64_t u_end = it.ulim < u0 + 4ull * n ? it.ulim : u0 + 4ull * n;
    uint64_t my_e = 0;
    if (u0 + t < u_end) my_e = gload(it.index + ((u0 + t) >> lg_spb));
    for (uint32_t q = 0; q < n; q++) {
        const uint64_t tile = it.first + q;
        const uint64_t ut = tile * 4;
        if (ut >= u_end) break;
        uint64_t carry = gload(it.refs + tile * 8 + j);
        const uint32_t cnt = u_end - ut < 4 ? (uint32_t)(u_end - ut) : 4u;
        uint64_t e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) e[u] = readlane64(my_e, 4 * q + u);
        delta_units_ss<true>(it.data, e, cnt, ut, mask, p, t, j, gl, carry, acc);
    }
}

// Layouts without a unit decoder: one value per lane through read_compr.
__device__ __attribute__((noinline, pure)) uint64_t item_generic(const ms_bentry *en, uint64_t first, uint32_t n,
                                                                 uint32_t t)
{
    const ms_colv c = en->c;
    const ms_pred p = en->p;
    uint64_t acc = 0;
    for (uint64_t k = first + t; k < first + n; k += 64) {
        const uint64_t v = read_compr(c, k);
        acc += pred_eval(p, v) ? v : 0ull;
    }
    return acc;
}

// delta_f layouts without a unit decoder (a step of 1, 2 or 4, or inner
// layouts other than 8-lane dynamic_vbp): n tiles of 2048 values from
// tile0 on, the differences read one by one (read_compr, as the
// column-level generic decode does).  Lane t takes values base + 64 k + t,
// k = 0 .. 31: all of one residue class r = t mod step, so each row's
// running sums are a scan over the lanes of the class (shuffles by
// multiples of the step) on top of the class's value before the row (the
// tile carry, then each row's last).
__device__ __attribute__((noinline, pure)) uint64_t item_delta_generic(const ms_bentry *en, uint64_t tile0, uint32_t n,
                                                                       uint32_t t)
{
    const ms_colv c = en->c;
    const ms_pred p = en->p;
    const uint64_t n_compr = parts_of(c).n_compr;
    const uint32_t ds = c.g.dstep, r = t & (ds - 1);
    uint64_t acc = 0;
    for (uint32_t q = 0; q < n; q++) {
        const uint64_t base = (tile0 + q) * MS_DELTA_TILE;
        if (base >= n_compr) break;
        uint64_t run = gload(c.refs + (tile0 + q) * 8 + r);
        for (uint32_t k = 0; k < MS_DELTA_TILE / 64; k++) {
            const uint64_t i = base + 64ull * k + t;
            if (base + 64ull * k >= n_compr) break;
            const bool in = i < n_compr;
            uint64_t d = in ? read_compr(c, i) : 0ull;
            for (uint32_t o = ds; o < 64; o <<= 1) {
                const uint64_t y = __shfl_up(d, o, MS_WAVE);
                if (t >= o) d += y;
            }
            const uint64_t x = run + d;
            acc += in && pred_eval(p, x) ? x : 0ull;
            run += __shfl(d, (int)(64 - ds + r), MS_WAVE);
        }
    }
    return acc;
}

// The loop body, instantiated by kernels that differ only in their
// register budget (waves per SIMD), a choice of the autotuner.  TIMED adds
// each item's wall-clock ticks (the constant-rate counter) to its view's
// slot of `ticks`.
template <bool TIMED = false, int NU = 4>
__device__ __forceinline__ void batch_loop(const ms_bentry *__restrict__ E, const ms_bitem *__restrict__ I,
                                           uint32_t n_items, uint64_t *__restrict__ part,
                                           unsigned long long *__restrict__ ticks = nullptr)
{
    const uint32_t w = blockIdx.x * WAVES + wave_index();
    const uint32_t W = gridDim.x * WAVES;
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & 7, gl = t >> 3;
    ms_bitem nxt;
    if (w < n_items) nxt = I[w];
    uint32_t run_entry = 0xffffffffu, run_slot = 0;
    uint64_t acc = 0;
    for (uint32_t i = w; i < n_items; i += W) {
        const ms_bitem it = nxt;
        if (i + W < n_items) nxt = I[i + W];   // the next item's descriptor, one item ahead
        if (it.entry != run_entry) {
            if (run_entry != 0xffffffffu) {
                const uint64_t s = wave_sum_u64(acc);
                if (t == 0) part[run_slot] = s;
            }
            run_entry = it.entry;
            run_slot = i;
            acc = 0;
        }
        const ms_bentry &en = E[it.entry];
        ms_pred p;
        p.lo = it.plo;
        p.hi = it.phi;
        p.neg = it.pflags & 1;
        p.empty = (it.pflags >> 1) & 1;
        const uint32_t kind = it.kind_n >> 28, n = it.kind_n & 0x0fffffffu;
        uint64_t t_start = 0;
        if (TIMED) t_start = wall_clock64();
        switch (kind) {
        case BI_STATIC: {
            // unit u of a static_vbp column: 64 bw bytes at u * 64 bw,
            // given to the indexed-layout routine as index entries
            const uint32_t bw = it.aux;
            for (uint32_t k = 0; k < n; k += NU) {
                const uint32_t cnt = n - k < (uint32_t)NU ? n - k : (uint32_t)NU;
                uint64_t e[4] = {0, 0, 0, 0}, r[4] = {0, 0, 0, 0};
#pragma unroll
                for (int u = 0; u < NU; u++) e[u] = (((it.first + k + u) * 64ull * bw) << 8) | bw;
                vbp_units_ss<false, false, NU>(it.data, e, r, cnt, 0, 0, p, en.tab, j, gl, acc);
            }
            break;
        }
        case BI_BLOCKS: item_blocks<false, NU>(en, it, p, n, t, j, gl, acc); break;
        case BI_FOR: item_blocks<true, NU>(en, it, p, n, t, j, gl, acc); break;
        case BI_DELTA: item_delta(it, p, n, t, t >> 3, t & 7, acc); break;   // lane-major threads
        case BI_RLE: acc += rle_runs_ss((const u64x2 *)it.data, it.first, it.first + n, p, t); break;
        case BI_RAW: {
            const uint64_t *x = (const uint64_t *)it.data;
            for (uint64_t k = it.first + t; k < it.first + n; k += 64) {
                const uint64_t v = gload(x + k);
                acc += pred_eval(p, v) ? v : 0ull;
            }
            break;
        }
        case BI_DELTA_GEN: acc += item_delta_generic(&E[it.entry], it.first, n, t); break;
        default:   // BI_GENERIC
            acc += item_generic(&E[it.entry], it.first, n, t);
            break;
        }
        if (TIMED) {
            __builtin_amdgcn_s_waitcnt(0);   // the item's loads have returned
            const uint64_t t_end = wall_clock64();
            if (t == 0 && it.tag != ~0ull) atomicAdd(ticks + it.tag, (unsigned long long)(t_end - t_start));
        }
    }
    if (run_entry != 0xffffffffu) {
        const uint64_t s = wave_sum_u64(acc);
        if (t == 0) part[run_slot] = s;
    }
}

__global__ __launch_bounds__(WG) void k_batch_select_sum(const ms_bentry *__restrict__ E,
                                                         const ms_bitem *__restrict__ I, uint32_t n_items,
                                                         uint64_t *__restrict__ part)
{
    batch_loop(E, I, n_items, part);
}

__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_batch_select_sum_w6(
    const ms_bentry *__restrict__ E, const ms_bitem *__restrict__ I, uint32_t n_items, uint64_t *__restrict__ part)
{
    batch_loop(E, I, n_items, part);
}

__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_select_sum_w8(
    const ms_bentry *__restrict__ E, const ms_bitem *__restrict__ I, uint32_t n_items, uint64_t *__restrict__ part)
{
    batch_loop(E, I, n_items, part);
}

// The same with two units' words in flight per wave instead of four (the
// tuner's unroll choice: fewer registers per wave, more waves).
__global__ __launch_bounds__(WG) void k_batch_select_sum_u2(const ms_bentry *__restrict__ E,
                                                            const ms_bitem *__restrict__ I, uint32_t n_items,
                                                            uint64_t *__restrict__ part)
{
    batch_loop<false, 2>(E, I, n_items, part);
}

__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_select_sum_w8_u2(
    const ms_bentry *__restrict__ E, const ms_bitem *__restrict__ I, uint32_t n_items, uint64_t *__restrict__ part)
{
    batch_loop<false, 2>(E, I, n_items, part);
}

// The timed run: per view, the wall-clock ticks its items took.
__global__ __launch_bounds__(WG) void k_batch_select_sum_timed(const ms_bentry *__restrict__ E,
                                                               const ms_bitem *__restrict__ I, uint32_t n_items,
                                                               uint64_t *__restrict__ part,
                                                               unsigned long long *__restrict__ ticks)
{
    batch_loop<true>(E, I, n_items, part, ticks);
}

// One workgroup (256 or 1024 threads) per entry: the sum of its partial
// slots [lo, min(hi, lo + W)).
__global__ __launch_bounds__(1024) void k_batch_finish(const ms_bentry *__restrict__ E,
                                                       const uint64_t *__restrict__ part, uint32_t W,
                                                       uint64_t *results)
{
    const uint64_t e = blockIdx.x;
    const uint64_t lo = E[e].item_lo, hi = E[e].item_hi < lo + W ? E[e].item_hi : lo + W;
    uint64_t acc = 0;
    for (uint64_t k = lo + threadIdx.x; k < hi; k += blockDim.x) acc += part[k];
    __shared__ uint64_t s_part[16];
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t sum = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; w++) sum += s_part[w];
        results[e] = sum;
    }
}

__global__ void k_batch_meta(const ms_bentry *E, uint64_t n, ms_col_meta *out)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const ms_dev_meta *m = E[e].c.meta;
    out[e].count = gload(&m->count);
    out[e].size_used = gload(&m->size_used);
    out[e].size_compr = gload(&m->size_compr);
    out[e].size_alloc = 0;
}

}  // namespace ms

using namespace ms;

// ------------------------------------------------------------------ host
struct ms_batch {
    std::vector<ms_bentry> E;
    std::vector<const ms_col *> cols;   // the entries' columns (blockwise_individual_f views)
    std::vector<ms_col_meta> meta;   // read at the first build
    std::vector<ms_bitem> items;
    ms_batch_config cfg;
    bool built = false, have_meta = false;
    uint32_t cus = 0;
    uint64_t bytes = 0;
    ms_bentry *dE = nullptr;
    ms_bitem *dI = nullptr;
    uint64_t *dP = nullptr;
    size_t capE = 0, capI = 0, capP = 0;
    uint32_t grid = 0;
    uint32_t finish_threads = WG;   // 1024 when an entry has thousands of partials
    std::vector<hipEvent_t> ev;     // kernel timing: a ring of (start, end) pairs
    uint32_t ev_next = 0, ev_used = 0;
    uint32_t holds = 0;             // live plans whose graph reads dE, dI and dP
    uint64_t ntags = 0;             // views of the queue (timed runs)
};

static void batch_release(void *p) { ((ms_batch *)p)->holds--; }

// A batch recorded into a plan keeps its queue until every such plan is
// freed: rebuilding or freeing it would leave the graph reading stale or
// released memory.
static int check_unheld(const ms_batch *b)
{
    if (b->holds) return ms_fail(MS_ERR_INVALID, "the batch is recorded in %u live plan(s); free them first", b->holds);
    return MS_OK;
}

static const ms_batch_config k_default_cfg = {4, 16, 2, 4096, 8192, 0, 0, 0, 0};

static bool cfg_ok(const ms_batch_config &c)
{
    return c.wg_per_cu >= 1 && c.wg_per_cu <= 16 && c.units_per_item >= 1 && c.units_per_item <= 64 &&
           c.tiles_per_item >= 1 && c.tiles_per_item <= 16 && c.runs_per_item >= 64 &&
           c.runs_per_item <= (1u << 27) && c.values_per_item >= 64 && c.values_per_item <= (1u << 27) &&
           (c.waves_per_simd == 0 || c.waves_per_simd == 6 || c.waves_per_simd == 8) &&
           c.units_per_item_for <= 64 && c.units_per_item_narrow <= 64 &&
           (c.unroll == 0 || c.unroll == 4 || (c.unroll == 2 && c.waves_per_simd != 6));
}

extern "C" int ms_batch_create(ms_batch **out)
{
    MS_TRY(ms_need_init());
    if (!out) return ms_fail(MS_ERR_INVALID, "null argument");
    ms_batch *b = new (std::nothrow) ms_batch();
    if (!b) return ms_fail(MS_ERR_NOMEM, "host allocation");
    b->cfg = k_default_cfg;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        delete b;
        return ms_fail(MS_ERR_HIP, "CU count");
    }
    b->cus = (uint32_t)cus;
    *out = b;
    return MS_OK;
}

static bool delta_batchable(const ms_geom &g)
{
    return g.leaf == MS_DYNAMIC_VBP && g.lanes == 8 && g.lg_spb != 0xff && g.lg_ipo != 0xff && g.dstep == 8;
}

extern "C" int ms_batch_enqueue(ms_batch *b, const ms_col *in, uint32_t cmp, uint64_t val, uint64_t *entry)
{
    if (!b || !in) return ms_fail(MS_ERR_INVALID, "null argument");
    MS_TRY(check_unheld(b));
    if (cmp > MS_CMP_GE) return ms_fail(MS_ERR_INVALID, "comparison %u", cmp);
    if (in->g.kind == MS_STATIC_VBP_PAD || in->g.kind == MS_K_WISE_NS || in->g.kind == MS_GROUP_SIMPLE)
        return ms_fail(MS_ERR_INVALID, "no batched decoder for format kind %u: decompress it first", in->g.kind);
    if (in->g.kind == MS_BLOCKWISE_INDIVIDUAL && (!in->bwi || in->bwi->lcount[BWK_GS]))
        return ms_fail(MS_ERR_INVALID, "no batched decoder for group_simple_f blocks: decompress the column first");
    ms_bentry e = {};
    MS_TRY(colv_of(in, e.c));
    e.p = make_pred(cmp, val);
    e.le = pred_le_form(e.p);
    e.tab = make_bwtab(e.p);
    if (entry) *entry = b->E.size();
    b->E.push_back(e);
    b->cols.push_back(in);
    b->built = false;
    b->have_meta = false;
    return MS_OK;
}

template <class T>
static int ensure(T **p, size_t &cap, size_t n)
{
    if (n <= cap && *p) return MS_OK;
    if (*p) MS_HIP(hipFree(*p));
    *p = nullptr;
    cap = 0;
    MS_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
    cap = std::max<size_t>(n, 1);
    return MS_OK;
}

// Cuts the entries into items, in entry order.
static void make_queue(ms_batch *b)
{
    const ms_batch_config &cf = b->cfg;
    // item sizes per format and bit-width class (0: the bit-packed default)
    const uint32_t upi_for = cf.units_per_item_for ? cf.units_per_item_for : cf.units_per_item;
    const uint32_t upi_narrow = cf.units_per_item_narrow ? cf.units_per_item_narrow : cf.units_per_item;
    b->items.clear();
    b->bytes = 0;
    b->ntags = 0;
    ms_bitem proto = {};
    auto add = [&](uint32_t e, uint32_t kind, uint64_t first, uint64_t n, uint64_t per) {
        for (uint64_t k = 0; k < n; k += per) {
            const uint64_t m = std::min<uint64_t>(per, n - k);
            ms_bitem it = proto;
            it.entry = e;
            it.kind_n = (kind << 28) | (uint32_t)m;
            it.first = first + k;
            b->items.push_back(it);
        }
    };
    for (uint32_t e = 0; e < b->E.size(); e++) {
        ms_bentry &en = b->E[e];
        const ms_col_meta &m = b->meta[e];
        const ms_geom &g = en.c.g;
        const ms_parts pt = ms_parts_of(g, m.count, m.size_used, m.size_compr);
        b->bytes += m.size_used;
        en.item_lo = b->items.size();
        proto = ms_bitem{};
        proto.data = en.c.data;
        proto.index = en.c.index;
        proto.refs = en.c.refs;
        proto.plo = en.p.lo;
        proto.phi = en.p.hi;
        proto.pflags = (uint8_t)((en.p.neg ? 1 : 0) | (en.p.empty ? 2 : 0) | (pred_le_form(en.p) ? 4 : 0));
        proto.lg_spb = g.lg_spb;
        proto.lg_ipo = g.lg_ipo;
        proto.aux = g.bw;
        proto.ulim = pt.n_compr / 512;   // BI_DELTA: units of the compressed part
        proto.tag = g.kind == MS_BLOCKWISE_INDIVIDUAL ? ~0ull : b->ntags++;
        switch (g.kind) {
        case MS_BLOCKWISE_INDIVIDUAL: {
            // every block's view as items of its own layout
            const ms_col *c = b->cols[e];
            const ms_bwi &h = *c->bwi;
            const uint64_t B = h.B, T = (B + MS_DELTA_TILE - 1) / MS_DELTA_TILE;
            for (uint64_t v = 0; v < h.nb; v++) {
                const ms_bwi_view &w = h.v[v];
                const ms_bitem keep = proto;
                proto.tag = b->ntags++;
                if (w.code == 0) {
                    proto.data = c->data + w.off;
                    add(e, BI_RAW, 0, B, cf.values_per_item);
                } else if (w.code <= 64) {
                    proto.data = c->data + w.off;
                    proto.aux = w.code;
                    add(e, BI_STATIC, 0, B / 512, w.code <= 8 ? upi_narrow : cf.units_per_item);
                } else if (w.code == MS_BWI_RLE) {
                    proto.data = c->data + w.off;
                    add(e, BI_RLE, 0, w.size / 16, cf.runs_per_item);
                } else {
                    const int k = w.code == MS_BWI_DYNAMIC_VBP ? BWK_DVBP
                                : w.code == MS_BWI_DELTA_DYNAMIC_VBP ? BWK_DELTA : BWK_FOR;
                    proto.data = c->data;   // absolute index entries
                    proto.index = c->index + h.idx_off[k] + (uint64_t)w.slot * (B / 512);
                    proto.lg_spb = 0;
                    proto.lg_ipo = 1;
                    proto.ulim = B / 512;
                    if (k == BWK_DELTA) {
                        proto.refs = c->refs + h.refs_off[k] + (uint64_t)w.slot * T * 8;
                        add(e, BI_DELTA, 0, (B / 512 + 3) / 4, cf.tiles_per_item);
                    } else {
                        proto.refs = k == BWK_FOR ? c->refs + h.refs_off[k] + (uint64_t)w.slot * (B / 1024) : nullptr;
                        add(e, k == BWK_FOR ? BI_FOR : BI_BLOCKS, 0, B / 512,
                            k == BWK_FOR ? upi_for : cf.units_per_item);
                    }
                }
                proto = keep;
            }
            break;
        }
        case MS_RLE:
            add(e, BI_RLE, 0, m.size_used / 16, cf.runs_per_item);
            break;
        case MS_UNCOMPR:
            break;
        case MS_STATIC_VBP:
            if (g.lanes == 8) add(e, BI_STATIC, 0, pt.n_compr / 512, g.bw <= 8 ? upi_narrow : cf.units_per_item);
            else add(e, BI_GENERIC, 0, pt.n_compr, cf.values_per_item);
            break;
        case MS_DYNAMIC_VBP:
        case MS_FOR:
        case MS_DELTA: {
            const uint64_t units = pt.n_compr / 512;
            const bool blocks = g.leaf == MS_DYNAMIC_VBP && g.lanes == 8 && g.lg_spb != 0xff &&
                                (g.kind == MS_DYNAMIC_VBP || g.lg_ipo != 0xff);
            if (g.kind == MS_DELTA && delta_batchable(g)) add(e, BI_DELTA, 0, (units + 3) / 4, cf.tiles_per_item);
            else if (g.kind == MS_DELTA)
                add(e, BI_DELTA_GEN, 0, (pt.n_compr + MS_DELTA_TILE - 1) / MS_DELTA_TILE, cf.tiles_per_item);
            else if (blocks) add(e, g.kind == MS_FOR ? BI_FOR : BI_BLOCKS, 0, units,
                                 g.kind == MS_FOR ? upi_for : cf.units_per_item);
            else add(e, BI_GENERIC, 0, pt.n_compr, cf.values_per_item);
            break;
        }
        default:
            break;
        }
        // the uncompressed values: the rest part, or the whole uncompr_f column
        proto.data = en.c.data + pt.rest_off;
        add(e, BI_RAW, 0, pt.n_rest, cf.values_per_item);
        en.item_hi = b->items.size();
    }
    b->grid = b->cus * cf.wg_per_cu;
    uint64_t most = 0;
    for (const ms_bentry &en : b->E) most = std::max<uint64_t>(most, en.item_hi - en.item_lo);
    most = std::min<uint64_t>(most, (uint64_t)b->grid * WAVES);
    b->finish_threads = most > 2048 ? 1024 : WG;
}

static int build(ms_batch *b, hipStream_t s)
{
    const size_t nE = b->E.size();
    MS_TRY(ensure(&b->dE, b->capE, nE));
    if (!b->have_meta) {
        MS_HIP(hipMemcpyAsync(b->dE, b->E.data(), nE * sizeof(ms_bentry), hipMemcpyHostToDevice, s));
        ms_col_meta *dm = nullptr;
        MS_HIP(hipMallocAsync((void **)&dm, std::max<size_t>(nE, 1) * sizeof(ms_col_meta), s));
        if (nE) hipLaunchKernelGGL(k_batch_meta, dim3((nE + 255) / 256), dim3(256), 0, s, b->dE, (uint64_t)nE, dm);
        MS_HIP(hipGetLastError());
        b->meta.resize(nE);
        MS_HIP(hipMemcpyAsync(b->meta.data(), dm, nE * sizeof(ms_col_meta), hipMemcpyDeviceToHost, s));
        MS_HIP(hipFreeAsync(dm, s));
        MS_HIP(hipStreamSynchronize(s));
        b->have_meta = true;
    }
    make_queue(b);
    if (b->items.size() >= (1ull << 32) - 1) return ms_fail(MS_ERR_CAPACITY, "too many work items");
    MS_TRY(ensure(&b->dI, b->capI, b->items.size()));
    MS_TRY(ensure(&b->dP, b->capP, b->items.size()));
    MS_HIP(hipMemcpyAsync(b->dE, b->E.data(), nE * sizeof(ms_bentry), hipMemcpyHostToDevice, s));
    MS_HIP(hipMemcpyAsync(b->dI, b->items.data(), b->items.size() * sizeof(ms_bitem), hipMemcpyHostToDevice, s));
    // host vectors must outlive the copies
    MS_HIP(hipStreamSynchronize(s));
    b->built = true;
    return MS_OK;
}

static int launch(ms_batch *b, uint64_t *d_results, hipStream_t s)
{
    const uint64_t nE = b->E.size();
    if (nE == 0) return MS_OK;
    const bool u2 = b->cfg.unroll == 2;
    auto kern = b->cfg.waves_per_simd == 8 ? (u2 ? k_batch_select_sum_w8_u2 : k_batch_select_sum_w8)
              : b->cfg.waves_per_simd == 6 ? k_batch_select_sum_w6
              : (u2 ? k_batch_select_sum_u2 : k_batch_select_sum);
    const uint32_t slots = plan_capturing() ? 0 : (uint32_t)b->ev.size() / 2;
    const uint32_t k = slots ? b->ev_next % slots : 0;
    if (slots) MS_HIP(hipEventRecord(b->ev[2 * k], s));
    hipLaunchKernelGGL(kern, dim3(b->grid), dim3(WG), 0, s, (const ms_bentry *)b->dE, (const ms_bitem *)b->dI,
                       (uint32_t)b->items.size(), b->dP);
    if (slots) {
        MS_HIP(hipEventRecord(b->ev[2 * k + 1], s));
        b->ev_next++;
        b->ev_used = std::min(b->ev_used + 1, slots);
    }
    MS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_finish, dim3((uint32_t)nE), dim3(b->finish_threads), 0, s, (const ms_bentry *)b->dE,
                       (const uint64_t *)b->dP, b->grid * (uint32_t)WAVES, d_results);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

extern "C" int ms_batch_run(ms_batch *b, uint64_t *d_results, void *stream)
{
    if (!b || (!d_results && !b->E.empty())) return ms_fail(MS_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (plan_capturing()) {
        // the graph keeps reading this queue: it must exist before the
        // capture (no allocation or synchronisation inside it)
        if (!b->built) return ms_fail(MS_ERR_INVALID, "a batch must run once before it is recorded into a plan");
        MS_TRY(plan_hold(batch_release, b));
        b->holds++;
        return launch(b, d_results, s);
    }
    if (!b->built) MS_TRY(build(b, s));
    return launch(b, d_results, s);
}

static int time_cfg(ms_batch *b, const ms_batch_config &c, uint64_t *d_res, uint32_t reps, hipStream_t s,
                    float *ms)
{
    b->cfg = c;
    MS_TRY(build(b, s));
    MS_TRY(launch(b, d_res, s));   // warm-up
    hipEvent_t e0, e1;
    MS_HIP(hipEventCreate(&e0));
    MS_HIP(hipEventCreate(&e1));
    MS_HIP(hipEventRecord(e0, s));
    for (uint32_t r = 0; r < reps; r++) MS_TRY(launch(b, d_res, s));
    MS_HIP(hipEventRecord(e1, s));
    MS_HIP(hipEventSynchronize(e1));
    MS_HIP(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MS_OK;
}

extern "C" int ms_batch_tune(ms_batch *b, uint64_t *d_results, uint32_t reps, ms_batch_config *best, void *stream)
{
    if (!b || !d_results) return ms_fail(MS_ERR_INVALID, "null argument");
    MS_TRY(check_unheld(b));
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "ms_batch_tune cannot be recorded into a plan");
    hipStream_t s = (hipStream_t)stream;
    if (reps == 0) reps = 5;
    ms_batch_config cur = b->cfg;
    float cur_ms;
    MS_TRY(time_cfg(b, cur, d_results, reps, s, &cur_ms));
    // coordinate descent over the fields, at most two sweeps
    static const uint32_t wg[] = {2, 3, 4, 6, 8};
    static const uint32_t upi[] = {2, 4, 8, 16, 32, 64};
    static const uint32_t tpi[] = {1, 2, 4, 8};
    static const uint32_t rpi[] = {1024, 4096, 16384};
    static const uint32_t vpi[] = {2048, 8192, 32768};
    static const uint32_t wps[] = {0, 6, 8};
    static const uint32_t unr[] = {4, 2};
    static const uint32_t upo[] = {0, 8, 16, 32, 64};   // for_f items; 0: as the other bit-packed layouts
    static const uint32_t upn[] = {0, 16, 32, 64};      // static_vbp of at most 8 bits
    struct axis {
        const uint32_t *vals;
        int nv;
        uint32_t ms_batch_config::*field;
    };
    static const axis axes[] = {{wps, 3, &ms_batch_config::waves_per_simd},
                                {unr, 2, &ms_batch_config::unroll},
                                {wg, 5, &ms_batch_config::wg_per_cu},
                                {upi, 6, &ms_batch_config::units_per_item},
                                {upo, 5, &ms_batch_config::units_per_item_for},
                                {upn, 4, &ms_batch_config::units_per_item_narrow},
                                {tpi, 4, &ms_batch_config::tiles_per_item},
                                {rpi, 3, &ms_batch_config::runs_per_item},
                                {vpi, 3, &ms_batch_config::values_per_item}};
    for (int sweep = 0; sweep < 2; sweep++) {
        bool changed = false;
        for (const axis &a : axes) {
            for (int k = 0; k < a.nv; k++) {
                ms_batch_config c = cur;
                uint32_t *field = &(c.*a.field);
                if (*field == a.vals[k] || (a.field == &ms_batch_config::unroll && *field == 0 && a.vals[k] == 4))
                    continue;
                *field = a.vals[k];
                if (!cfg_ok(c)) continue;
                float t;
                MS_TRY(time_cfg(b, c, d_results, reps, s, &t));
                if (t < cur_ms * 0.99f) {
                    cur = c;
                    cur_ms = t;
                    changed = true;
                }
            }
        }
        if (!changed) break;
    }
    b->cfg = cur;
    MS_TRY(build(b, s));
    if (best) *best = cur;
    return MS_OK;
}

extern "C" int ms_batch_get_config(const ms_batch *b, ms_batch_config *cfg)
{
    if (!b || !cfg) return ms_fail(MS_ERR_INVALID, "null argument");
    *cfg = b->cfg;
    return MS_OK;
}

extern "C" int ms_batch_set_config(ms_batch *b, const ms_batch_config *cfg)
{
    if (!b || !cfg) return ms_fail(MS_ERR_INVALID, "null argument");
    MS_TRY(check_unheld(b));
    if (!cfg_ok(*cfg)) return ms_fail(MS_ERR_INVALID, "batch configuration out of range");
    b->cfg = *cfg;
    b->built = false;
    return MS_OK;
}

extern "C" int ms_batch_stats(const ms_batch *b, uint64_t *entries, uint64_t *items, uint64_t *bytes)
{
    if (!b) return ms_fail(MS_ERR_INVALID, "null argument");
    if (entries) *entries = b->E.size();
    if (items) *items = b->items.size();
    if (bytes) *bytes = b->bytes;
    return MS_OK;
}

static void drop_events(ms_batch *b)
{
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    b->ev.clear();
    b->ev_next = b->ev_used = 0;
}

extern "C" int ms_batch_timing(ms_batch *b, uint32_t slots)
{
    if (!b) return ms_fail(MS_ERR_INVALID, "null argument");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "ms_batch_timing cannot be recorded into a plan");
    if (slots > 4096) return ms_fail(MS_ERR_INVALID, "at most 4096 timing slots");
    MS_HIP(hipDeviceSynchronize());
    drop_events(b);
    for (uint32_t i = 0; i < 2 * slots; i++) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            drop_events(b);
            return ms_fail(MS_ERR_HIP, "hipEventCreate");
        }
        b->ev.push_back(e);
    }
    return MS_OK;
}

extern "C" int ms_batch_kernel_times(ms_batch *b, float *ms, uint32_t *n)
{
    if (!b || !ms || !n) return ms_fail(MS_ERR_INVALID, "null argument");
    const uint32_t slots = (uint32_t)b->ev.size() / 2;
    const uint32_t cnt = std::min(*n, b->ev_used);
    for (uint32_t i = 0; i < cnt; i++) {
        const uint32_t k = (b->ev_next - cnt + i) % slots;
        MS_HIP(hipEventSynchronize(b->ev[2 * k + 1]));
        MS_HIP(hipEventElapsedTime(&ms[i], b->ev[2 * k], b->ev[2 * k + 1]));
    }
    *n = cnt;
    return MS_OK;
}

extern "C" int ms_batch_view_times(ms_batch *b, uint64_t *d_results, uint32_t reps, double *ns, uint64_t cap,
                                   uint64_t *nviews, void *stream)
{
    if (!b || !d_results || !nviews || reps == 0) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "timed runs synchronise: not in a plan");
    MS_TRY(check_unheld(b));
    hipStream_t s = (hipStream_t)stream;
    if (!b->built) MS_TRY(build(b, s));
    *nviews = b->ntags;
    if (b->E.empty() || b->items.empty()) return MS_OK;
    int dev = 0, khz = 0;
    MS_HIP(hipGetDevice(&dev));
    MS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return ms_fail(MS_ERR_HIP, "no wall clock rate");
    unsigned long long *ticks = nullptr;
    MS_HIP(hipMallocAsync((void **)&ticks, std::max<uint64_t>(b->ntags, 1) * 8, s));
    int rc = launch_fill_u64((uint64_t *)ticks, std::max<uint64_t>(b->ntags, 1), 0, s);
    for (uint32_t r = 0; r < reps && !rc; r++) {
        hipLaunchKernelGGL(k_batch_select_sum_timed, dim3(b->grid), dim3(WG), 0, s, (const ms_bentry *)b->dE,
                           (const ms_bitem *)b->dI, (uint32_t)b->items.size(), b->dP, ticks);
        if (hipGetLastError() != hipSuccess) rc = ms_fail(MS_ERR_HIP, "timed launch");
    }
    if (!rc) {
        hipLaunchKernelGGL(k_batch_finish, dim3((uint32_t)b->E.size()), dim3(b->finish_threads), 0, s,
                           (const ms_bentry *)b->dE, (const uint64_t *)b->dP, b->grid * (uint32_t)WAVES, d_results);
        if (hipGetLastError() != hipSuccess) rc = ms_fail(MS_ERR_HIP, "finish launch");
    }
    std::vector<unsigned long long> h(b->ntags);
    if (!rc && b->ntags &&
        hipMemcpyAsync(h.data(), ticks, b->ntags * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = ms_fail(MS_ERR_HIP, "copy");
    (void)hipFreeAsync(ticks, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = ms_fail(MS_ERR_HIP, "synchronise");
    if (rc) return rc;
    for (uint64_t v = 0; v < b->ntags && v < cap && ns; v++) ns[v] = (double)h[v] / reps * 1e6 / khz;
    return MS_OK;
}

extern "C" int ms_batch_free(ms_batch *b)
{
    if (!b) return MS_OK;
    MS_TRY(check_unheld(b));
    (void)hipDeviceSynchronize();
    drop_events(b);
    if (b->dE) (void)hipFree(b->dE);
    if (b->dI) (void)hipFree(b->dI);
    if (b->dP) (void)hipFree(b->dP);
    delete b;
    return MS_OK;
}
