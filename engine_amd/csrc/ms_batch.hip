// ms_batch.hip -- fused select + agg_sum over many columns of uneven length
// and mixed formats in one persistent launch.
//
// The host cuts every enqueued column into work items (a run of 512-value
// units, of 2048-value delta tiles, of rle runs, or of raw values) and
// splits the item list into one contiguous range per wave so that every
// wave gets the same modelled cost (compressed bytes plus a fixed cost per
// unit).  The persistent kernel walks its wave's range of that device work
// queue; each item's select + sum goes to its own partial slot (no atomics),
// and a second kernel folds the partials of each entry.  The queue shape
// (workgroups per CU, item sizes per layout, the unit cost of the model) is
// a run-time configuration that ms_batch_tune picks by timing candidates.
#include <algorithm>
#include <vector>

#include "ms_units.hpp"

namespace ms {

enum { BI_STATIC = 0, BI_BLOCKS = 1, BI_FOR = 2, BI_DELTA = 3, BI_RLE = 4, BI_RAW = 5, BI_GENERIC = 6 };

struct ms_bentry {
    ms_colv c;
    ms_pred p;
    uint64_t item_lo, item_hi;   // the entry's items in the queue
};

struct ms_bitem {
    uint32_t entry;
    uint32_t kind_n;   // kind << 28 | number of units / tiles / runs / values
    uint64_t first;    // first unit / tile / run / value of the item
};

// Per-unit predicate on codes for x = ref + code, bw <= 32, no wrap.
__device__ __forceinline__ uint32_t unit_mode(const ms_pred &p, uint32_t bw, uint64_t ref, uint32_t &lo,
                                              uint32_t &span)
{
    const ms_cpred cp = code_pred(p, ref, (1ull << bw) - 1ull);
    lo = (uint32_t)cp.clo;
    span = (uint32_t)cp.cspan;
    return cp.mode == PM_NONE ? SM_NONE : cp.mode == PM_ALL ? SM_ALL : cp.mode == PM_OUT ? SM_OUT
         : cp.clo == 0 ? SM_LE : SM_IN;
}

// Value-by-value select + sum of a thread's 8 codes (wide or wrapping):
// rare, kept out of line so it does not set the kernel's register budget
// (and declared pure, so the kernel's uniform loads stay scalar).
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

// One unit of an 8-lane bit-packed layout: `base` is the unit's block,
// g the thread's code group inside it.
template <bool IS_FOR>
__device__ __forceinline__ void vbp_unit(const ms_pred &p, const uint64_t *base, uint32_t j, uint32_t g,
                                         uint32_t bw, uint64_t ref, uint64_t &acc)
{
    if (bw > 32 || (IS_FOR && ref + ((1ull << bw) - 1ull) < ref)) {
        acc += slow_codes(p, base, j, g, bw, ref);
        return;
    }
    uint32_t lo, span;
    const uint32_t mode = unit_mode(p, bw, ref, lo, span);
    if (mode == SM_NONE) return;
    ss_fast_dispatch<IS_FOR>(bw, base, j, g, mode, lo, span, ref, acc);
}

template <bool IS_FOR>
__device__ __forceinline__ void item_blocks(const ms_colv &c, const ms_pred &p, uint64_t u0, uint32_t n, uint32_t j,
                                            uint32_t gl, uint64_t &acc)
{
    const uint32_t lg_spb = c.g.lg_spb, lg_ipo = c.g.lg_ipo;
    const uint64_t mask = (1ull << lg_spb) - 1ull;
    uint64_t e_next = uniform64(c.index[u0 >> lg_spb]);
    uint64_t r_next = IS_FOR ? uniform64(c.refs[(u0 >> lg_spb) >> lg_ipo]) : 0ull;
    for (uint64_t u = u0; u < u0 + n; u++) {
        const uint64_t e = e_next;
        const uint64_t ref = r_next;
        if (u + 1 < u0 + n) {
            e_next = uniform64(c.index[(u + 1) >> lg_spb]);
            if (IS_FOR) r_next = uniform64(c.refs[((u + 1) >> lg_spb) >> lg_ipo]);
        }
        vbp_unit<IS_FOR>(p, (const uint64_t *)(c.data + (e >> 8)), j, (uint32_t)(u & mask) * 8 + gl,
                         (uint32_t)(e & 0xff), ref, acc);
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

__global__ __launch_bounds__(WG) void k_batch_select_sum(const ms_bentry *__restrict__ E,
                                                         const ms_bitem *__restrict__ I,
                                                         const uint32_t *__restrict__ wave_lo, uint64_t *__restrict__ part)
{
    const uint32_t w = blockIdx.x * WAVES + wave_index();
    const uint32_t i0 = __builtin_amdgcn_readfirstlane(wave_lo[w]);
    const uint32_t i1 = __builtin_amdgcn_readfirstlane(wave_lo[w + 1]);
    const uint32_t t = threadIdx.x & 63;
    const uint32_t j = t & 7, gl = t >> 3;
    for (uint32_t i = i0; i < i1; i++) {
        const ms_bitem it = I[i];
        const ms_bentry &en = E[it.entry];
        const ms_colv c = en.c;
        const ms_pred p = en.p;
        const uint32_t kind = it.kind_n >> 28, n = it.kind_n & 0x0fffffffu;
        uint64_t acc = 0;
        switch (kind) {
        case BI_STATIC: {
            const uint32_t bw = c.g.bw;
            for (uint64_t u = it.first; u < it.first + n; u++) {
                // the code group is loop invariant here; hiding that keeps
                // the compiler from hoisting per-width offsets of all 32
                // cases into registers
                uint32_t g = gl;
                asm volatile("" : "+v"(g));
                vbp_unit<false>(p, (const uint64_t *)c.data + u * 8 * bw, j, g, bw, 0, acc);
            }
            break;
        }
        case BI_BLOCKS: item_blocks<false>(c, p, it.first, n, j, gl, acc); break;
        case BI_FOR: item_blocks<true>(c, p, it.first, n, j, gl, acc); break;
        case BI_DELTA: {
            const ms_parts pt = parts_of(c);
            for (uint64_t tile = it.first; tile < it.first + n; tile++)
                delta_tile_ss(c, p, tile, pt.n_compr >> 9, t, j, gl, acc);
            break;
        }
        case BI_RLE: {
            const u64x2 *r = (const u64x2 *)c.data;
            for (uint64_t k = it.first + t; k < it.first + n; k += 64) {
                const u64x2 vl = r[k];
                if (pred_eval(p, vl.x)) acc += vl.x * vl.y;
            }
            break;
        }
        case BI_RAW: {
            const ms_parts pt = parts_of(c);
            const uint64_t *x = (const uint64_t *)(c.data + pt.rest_off);
            for (uint64_t k = it.first + t; k < it.first + n; k += 64) {
                const uint64_t v = x[k];
                acc += pred_eval(p, v) ? v : 0ull;
            }
            break;
        }
        default:   // BI_GENERIC
            acc = item_generic(&E[it.entry], it.first, n, t);
            break;
        }
        acc = wave_sum_u64(acc);
        if (t == 0) part[i] = acc;
    }
}

// One wave per entry: the sum of its items' partials.
__global__ __launch_bounds__(WG) void k_batch_finish(const ms_bentry *__restrict__ E, uint64_t n_entries,
                                                     const uint64_t *__restrict__ part, uint64_t *results)
{
    const uint64_t e = (uint64_t)blockIdx.x * WAVES + wave_index();
    if (e >= n_entries) return;
    const uint32_t t = threadIdx.x & 63;
    uint64_t acc = 0;
    for (uint64_t k = E[e].item_lo + t; k < E[e].item_hi; k += 64) acc += part[k];
    acc = wave_sum_u64(acc);
    if (t == 0) results[e] = acc;
}

__global__ void k_batch_meta(const ms_bentry *E, uint64_t n, ms_col_meta *out)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const ms_dev_meta *m = E[e].c.meta;
    out[e].count = m->count;
    out[e].size_used = m->size_used;
    out[e].size_compr = m->size_compr;
    out[e].size_alloc = 0;
}

}  // namespace ms

using namespace ms;

// ------------------------------------------------------------------ host
struct ms_batch {
    std::vector<ms_bentry> E;
    std::vector<ms_col_meta> meta;   // read at the first build
    std::vector<ms_bitem> items;
    std::vector<uint32_t> wave_lo;
    ms_batch_config cfg;
    bool built = false, have_meta = false;
    uint32_t cus = 0;
    uint64_t bytes = 0;
    ms_bentry *dE = nullptr;
    ms_bitem *dI = nullptr;
    uint32_t *dW = nullptr;
    uint64_t *dP = nullptr;
    size_t capE = 0, capI = 0, capW = 0, capP = 0;
    uint32_t grid = 0;
};

static const ms_batch_config k_default_cfg = {4, 8, 2, 4096, 8192, 256};

static bool cfg_ok(const ms_batch_config &c)
{
    return c.wg_per_cu >= 1 && c.wg_per_cu <= 16 && c.units_per_item >= 1 && c.units_per_item <= (1u << 20) &&
           c.tiles_per_item >= 1 && c.tiles_per_item <= (1u << 20) && c.runs_per_item >= 64 &&
           c.runs_per_item <= (1u << 27) && c.values_per_item >= 64 && c.values_per_item <= (1u << 27);
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
    if (cmp > MS_CMP_GE) return ms_fail(MS_ERR_INVALID, "comparison %u", cmp);
    if (in->g.kind == MS_DELTA && !delta_batchable(in->g))
        return ms_fail(MS_ERR_INVALID, "batched delta_f needs step 8 over an 8-lane dynamic_vbp inner format");
    ms_bentry e = {};
    MS_TRY(colv_of(in, e.c));
    e.p = make_pred(cmp, val);
    if (entry) *entry = b->E.size();
    b->E.push_back(e);
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

// Cuts the entries into items and the item list into per-wave ranges.
static void make_queue(ms_batch *b)
{
    const ms_batch_config &cf = b->cfg;
    b->items.clear();
    std::vector<double> cost;
    b->bytes = 0;
    auto add = [&](uint32_t e, uint32_t kind, uint64_t first, uint64_t n, uint64_t per, double unit_bytes,
                   double fixed) {
        for (uint64_t k = 0; k < n; k += per) {
            const uint64_t m = std::min<uint64_t>(per, n - k);
            b->items.push_back(ms_bitem{e, (kind << 28) | (uint32_t)m, first + k});
            cost.push_back(m * (unit_bytes + fixed) + 256.0);
        }
    };
    for (uint32_t e = 0; e < b->E.size(); e++) {
        ms_bentry &en = b->E[e];
        const ms_col_meta &m = b->meta[e];
        const ms_geom &g = en.c.g;
        const ms_parts pt = ms_parts_of(g, m.count, m.size_used, m.size_compr);
        b->bytes += m.size_used;
        en.item_lo = b->items.size();
        const double uc = cf.unit_cost;
        switch (g.kind) {
        case MS_RLE:
            add(e, BI_RLE, 0, m.size_used / 16, cf.runs_per_item, 16.0, 0.0);
            break;
        case MS_UNCOMPR:
            break;
        case MS_STATIC_VBP:
            if (g.lanes == 8) add(e, BI_STATIC, 0, pt.n_compr / 512, cf.units_per_item, 64.0 * g.bw, uc);
            else add(e, BI_GENERIC, 0, pt.n_compr, cf.values_per_item, 32.0, 0.0);
            break;
        case MS_DYNAMIC_VBP:
        case MS_FOR:
        case MS_DELTA: {
            const uint64_t units = pt.n_compr / 512;
            const double ub = units ? (double)m.size_compr / units : 0.0;
            const bool blocks = g.leaf == MS_DYNAMIC_VBP && g.lanes == 8 && g.lg_spb != 0xff &&
                                (g.kind == MS_DYNAMIC_VBP || g.lg_ipo != 0xff);
            if (g.kind == MS_DELTA) add(e, BI_DELTA, 0, (units + 3) / 4, cf.tiles_per_item, 4 * ub, 6 * uc);
            else if (blocks) add(e, g.kind == MS_FOR ? BI_FOR : BI_BLOCKS, 0, units, cf.units_per_item, ub, uc);
            else add(e, BI_GENERIC, 0, pt.n_compr, cf.values_per_item, 32.0, 0.0);
            break;
        }
        default:
            break;
        }
        // the uncompressed values: the rest part, or the whole uncompr_f column
        add(e, BI_RAW, 0, pt.n_rest, cf.values_per_item, 8.0, 0.0);
        en.item_hi = b->items.size();
    }
    // equal-cost contiguous ranges, one per wave
    b->grid = b->cus * cf.wg_per_cu;
    const uint64_t W = (uint64_t)b->grid * WAVES;
    double total = 0;
    for (double c : cost) total += c;
    b->wave_lo.assign(W + 1, (uint32_t)b->items.size());
    double run = 0;
    uint64_t w = 0;
    b->wave_lo[0] = 0;
    for (size_t i = 0; i < cost.size() && total > 0; i++) {
        // item i starts in the range of the wave its cost midpoint falls in
        const double mid = run + cost[i] / 2;
        const uint64_t owner = std::min<uint64_t>(W - 1, (uint64_t)(mid / total * W));
        while (w < owner) b->wave_lo[++w] = (uint32_t)i;
        run += cost[i];
    }
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
    MS_TRY(ensure(&b->dW, b->capW, b->wave_lo.size()));
    MS_HIP(hipMemcpyAsync(b->dE, b->E.data(), nE * sizeof(ms_bentry), hipMemcpyHostToDevice, s));
    MS_HIP(hipMemcpyAsync(b->dI, b->items.data(), b->items.size() * sizeof(ms_bitem), hipMemcpyHostToDevice, s));
    MS_HIP(hipMemcpyAsync(b->dW, b->wave_lo.data(), b->wave_lo.size() * 4, hipMemcpyHostToDevice, s));
    // host vectors must outlive the copies
    MS_HIP(hipStreamSynchronize(s));
    b->built = true;
    return MS_OK;
}

static int launch(ms_batch *b, uint64_t *d_results, hipStream_t s)
{
    const uint64_t nE = b->E.size();
    if (nE == 0) return MS_OK;
    hipLaunchKernelGGL(k_batch_select_sum, dim3(b->grid), dim3(WG), 0, s, (const ms_bentry *)b->dE,
                       (const ms_bitem *)b->dI, (const uint32_t *)b->dW, b->dP);
    MS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_batch_finish, dim3((nE + WAVES - 1) / WAVES), dim3(WG), 0, s, (const ms_bentry *)b->dE, nE,
                       (const uint64_t *)b->dP, d_results);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

extern "C" int ms_batch_run(ms_batch *b, uint64_t *d_results, void *stream)
{
    if (!b || (!d_results && !b->E.empty())) return ms_fail(MS_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
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
    hipStream_t s = (hipStream_t)stream;
    if (reps == 0) reps = 5;
    ms_batch_config cur = b->cfg;
    float cur_ms;
    MS_TRY(time_cfg(b, cur, d_results, reps, s, &cur_ms));
    // coordinate descent over the fields, two sweeps
    static const uint32_t wg[] = {1, 2, 3, 4, 6, 8};
    static const uint32_t upi[] = {1, 2, 4, 8, 16, 32};
    static const uint32_t tpi[] = {1, 2, 4, 8};
    static const uint32_t rpi[] = {1024, 4096, 16384};
    static const uint32_t vpi[] = {2048, 8192, 32768};
    static const uint32_t uco[] = {0, 256, 1024};
    for (int sweep = 0; sweep < 2; sweep++) {
        bool changed = false;
        for (int f = 0; f < 6; f++) {
            const uint32_t *vals = f == 0 ? wg : f == 1 ? upi : f == 2 ? tpi : f == 3 ? rpi : f == 4 ? vpi : uco;
            const int nv = f == 0 ? 6 : f == 1 ? 6 : f == 2 ? 4 : 3;
            for (int k = 0; k < nv; k++) {
                ms_batch_config c = cur;
                uint32_t *field = f == 0 ? &c.wg_per_cu : f == 1 ? &c.units_per_item : f == 2 ? &c.tiles_per_item
                                : f == 3 ? &c.runs_per_item : f == 4 ? &c.values_per_item : &c.unit_cost;
                if (*field == vals[k]) continue;
                *field = vals[k];
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

extern "C" int ms_batch_free(ms_batch *b)
{
    if (!b) return MS_OK;
    (void)hipDeviceSynchronize();
    if (b->dE) (void)hipFree(b->dE);
    if (b->dI) (void)hipFree(b->dI);
    if (b->dW) (void)hipFree(b->dW);
    if (b->dP) (void)hipFree(b->dP);
    delete b;
    return MS_OK;
}
