// ms_abi.hip -- the C-ABI of include/morphstore_hip.h: column management
// and operator orchestration.  Entry points never throw; errors are status
// codes with a per-thread message (ms_last_error).
#include <mutex>
#include <unordered_set>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "ms_kernels.h"

static thread_local std::string g_err;
static int g_device = -1;

// Device memory of the entry points that allocate nothing
// (ms_select_sum_noalloc): SS_MAX_PARTS partial sums followed by a
// temporary of NOALLOC_TMP_VALUES decoded deltas.  Each device holds its
// own copy for the life of the process; host code reaches it only through
// the address hipGetSymbolAddress returned for that device at ms_init.
static constexpr uint64_t NOALLOC_TMP_VALUES = 1ull << 20;
static constexpr uint64_t NOALLOC_WORDS = ms::SS_MAX_PARTS + NOALLOC_TMP_VALUES;
static __device__ __align__(256) uint64_t scratch[NOALLOC_WORDS];
static constexpr int MAX_DEVICES = 64;
static uint64_t *g_scratch_ptr[MAX_DEVICES];   // per device, set by ms_init

int ms_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// ------------------------------------------------------------------ plans
// While a plan records (ms_plan_begin .. ms_plan_end) every device
// allocation of the operators on the calling thread comes from the plan's
// arena, so the captured hipGraph holds kernels and copies only and can be
// replayed.  A plan created with arena_bytes == 0 first runs in measuring
// mode: the operators run eagerly and the plan records the bytes they
// allocate; the next ms_plan_begin captures with an arena of that size.
struct ms_plan {
    int mode;              // 0: measure, 1: capture
    uint8_t *arena;
    uint64_t arena_bytes;
    uint64_t used;
    uint64_t measured;
    hipStream_t stream;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool recording;
    std::vector<void *> deferred;   // frees of eager memory requested while capturing
    std::vector<std::pair<void (*)(void *), void *>> holds;   // objects the graph reads
};

static thread_local ms_plan *g_rec = nullptr;

static uint64_t round256(uint64_t x) { return (x + 255) / 256 * 256; }

namespace ms {

bool plan_capturing() { return g_rec && g_rec->recording && g_rec->mode == 1; }

int plan_hold(void (*release)(void *), void *obj)
{
    if (!plan_capturing()) return ms_fail(MS_ERR_INVALID, "no plan is capturing");
    g_rec->holds.emplace_back(release, obj);
    return MS_OK;
}

void *dev_alloc(size_t bytes, hipStream_t s)
{
    bytes = round256(bytes ? bytes : 64);
    if (g_rec && g_rec->recording && g_rec->mode == 1) {
        if (g_rec->used + bytes > g_rec->arena_bytes) return nullptr;
        void *p = g_rec->arena + g_rec->used;
        g_rec->used += bytes;
        return p;
    }
    void *p = nullptr;
    if (hipMallocAsync(&p, bytes, s) != hipSuccess) return nullptr;
    if (g_rec && g_rec->recording && g_rec->mode == 0) g_rec->measured += bytes;
    return p;
}

static bool in_arena(const void *p)
{
    return g_rec && g_rec->arena && (const uint8_t *)p >= g_rec->arena &&
           (const uint8_t *)p < g_rec->arena + g_rec->arena_bytes;
}

void dev_free(void *p, hipStream_t s)
{
    if (!p || in_arena(p)) return;
    (void)hipFreeAsync(p, s);
}

void *temps::get(size_t bytes)
{
    if (n >= (int)(sizeof ptrs / sizeof ptrs[0])) return nullptr;
    void *p = dev_alloc(bytes, s);
    if (!p) return nullptr;
    ptrs[n++] = p;
    return p;
}

temps::~temps()
{
    for (int i = 0; i < n; i++) dev_free(ptrs[i], s);
}

int colv_of(const ms_col *c, ms_colv &v)
{
    if (!c) return ms_fail(MS_ERR_INVALID, "null column");
    v.data = c->data;
    v.meta = c->meta;
    v.index = c->index;
    v.refs = c->refs;
    v.g = c->g;
    return MS_OK;
}

}  // namespace ms

using namespace ms;

// ------------------------------------------------------------------ formats
static bool valid_leaf(const ms_format_desc &f)
{
    switch (f.kind) {
    case MS_UNCOMPR: return true;
    case MS_STATIC_VBP: return f.bw >= 1 && f.bw <= 64 && f.step >= 1 && f.step <= 64;
    case MS_DYNAMIC_VBP:
        return f.step >= 1 && f.step <= 64 && f.block > 0 && f.page > 0 && f.block % (64u * f.step) == 0 &&
               f.page % (8u * f.step) == 0;
    default: return false;
    }
}

static bool valid_format(const ms_format_desc *fp)
{
    if (!fp) return false;
    const ms_format_desc &f = *fp;
    switch (f.kind) {
    case MS_UNCOMPR:
    case MS_STATIC_VBP:
    case MS_DYNAMIC_VBP: return valid_leaf(f);
    case MS_RLE: return true;
    case MS_STATIC_VBP_PAD: return f.bw >= 1 && f.bw <= 64 && f.step >= 1 && f.step <= 64;
    case MS_K_WISE_NS: return f.block == 2;   // the reference implements 2-wise null suppression only
    case MS_GROUP_SIMPLE: return f.step >= 1 && f.step <= 64;
    case MS_BLOCKWISE_INDIVIDUAL:   // blockwise_individual.h:125-139: a multiple of every default format's block
        return f.block > 0 && f.block % 1024 == 0 && f.block <= (1u << 20) && f.step == 8 &&
               f.bw <= MS_BWI_SEL_STATIC && (f.bw != MS_BWI_SEL_STATIC || bwi_code_ok(f.page));
    case MS_DELTA:
    case MS_FOR: {
        ms_format_desc in = {};
        in.kind = f.inner_kind; in.bw = f.inner_bw; in.block = f.inner_block; in.page = f.inner_page;
        in.step = f.inner_step;
        if (!valid_leaf(in) || f.block == 0) return false;
        const uint64_t ibs = in.kind == MS_UNCOMPR ? 1 : in.kind == MS_STATIC_VBP ? 64ull * in.step : in.block;
        if (f.block % ibs) return false;
        if (f.kind == MS_DELTA)
            return f.step >= 1 && f.step <= 8 && f.block % f.step == 0 &&
                   (in.kind == MS_UNCOMPR || in.step == f.step) && (8 % f.step == 0);
        const uint32_t E = f.step ? f.step : (in.kind != MS_UNCOMPR ? in.step : 1);
        return f.page > 0 && f.page % E == 0 && f.block % E == 0 && (in.kind == MS_UNCOMPR || in.step == E) &&
               (64 % E == 0);
    }
    default: return false;
    }
}

static uint64_t size_max(const ms_format_desc &f, uint64_t n)
{
    const ms_geom g = ms_make_geom(f);
    switch (f.kind) {
    case MS_UNCOMPR: return 8 * n;
    case MS_STATIC_VBP: return n * f.bw / 8;
    case MS_DYNAMIC_VBP: return 8 * n + (n / f.block) * f.page;
    case MS_DELTA: return ms_leaf_size_max(f.inner_kind, f.inner_bw, f.inner_block, f.inner_page, n);
    case MS_FOR: {
        const uint64_t pages = (n + (uint64_t)f.page * f.block - 1) / ((uint64_t)f.page * f.block);
        return pages * 8ull * f.page +
               (n / f.block) * ms_leaf_size_max(f.inner_kind, f.inner_bw, f.inner_block, f.inner_page, f.block);
    }
    case MS_RLE: return 16 * n;
    case MS_STATIC_VBP_PAD: return ms_pad_size_max(f.bw, f.step, n);
    case MS_K_WISE_NS: return ms_kwns_size_max(f.block, n);
    case MS_GROUP_SIMPLE: return ms_gs_size_max(f.step, n);
    case MS_BLOCKWISE_INDIVIDUAL: return ms_bwi_size_max(f.block, n);
    default: (void)g; return 0;
    }
}

// Formats of the extra codecs (ms_xfmt.hip): column-level morph and
// decompression only.
static bool xfmt(uint32_t kind) { return kind == MS_STATIC_VBP_PAD || kind == MS_K_WISE_NS || kind == MS_GROUP_SIMPLE; }

static uint64_t size_max_any_len(const ms_format_desc &f, uint64_t n)
{
    if (f.kind == MS_UNCOMPR) return 8 * n;
    if (f.kind == MS_RLE) return 16 * n;
    const uint64_t bs = ms_make_geom(f).bs;
    return size_max(f, (n + bs - 1) / bs * bs + bs) + 8 * (n % bs + bs) + 63;
}

extern "C" int ms_size_max_byte_any_len(const ms_format_desc *f, uint64_t count, uint64_t *out)
{
    if (!valid_format(f) || !out) return ms_fail(MS_ERR_INVALID, "invalid format");
    *out = size_max_any_len(*f, count);
    return MS_OK;
}

extern "C" int ms_size_max_byte(const ms_format_desc *f, uint64_t count, uint64_t *out)
{
    if (!valid_format(f) || !out) return ms_fail(MS_ERR_INVALID, "invalid format");
    *out = size_max(*f, count);
    return MS_OK;
}

extern "C" int ms_block_size(const ms_format_desc *f, uint64_t *out)
{
    if (!valid_format(f) || !out || f->kind == MS_RLE) return ms_fail(MS_ERR_INVALID, "format has no block size");
    *out = ms_make_geom(*f).bs;
    return MS_OK;
}

extern "C" const char *ms_last_error(void) { return g_err.c_str(); }
extern "C" const char *ms_version(void) { return "morphstore-hip 0.1 (gfx950)"; }

extern "C" int ms_init(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
        return ms_fail(MS_ERR_NODEVICE, "no HIP device %d", device);
    hipDeviceProp_t prop;
    MS_HIP(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return ms_fail(MS_ERR_NODEVICE, "device %d is %s, not gfx950", device, prop.gcnArchName);
    MS_HIP(hipSetDevice(device));
    hipMemPool_t pool;
    MS_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t thr = ~0ull;
    MS_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    if (device >= MAX_DEVICES) return ms_fail(MS_ERR_NODEVICE, "device %d beyond the %d supported", device, MAX_DEVICES);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        MS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ms_fail(MS_ERR_HIP, "no device address for the no-allocation scratch array");
        g_scratch_ptr[device] = (uint64_t *)p;
    }
    g_device = device;
    return MS_OK;
}

int ms_need_init()
{
    if (g_device < 0) return ms_fail(MS_ERR_NODEVICE, "ms_init was not called");
    return MS_OK;
}

// ------------------------------------------------------------------ columns
static uint64_t idx_entries(const ms_geom &g, uint64_t count)
{
    if (g.kind == MS_DYNAMIC_VBP || g.kind == MS_DELTA || g.kind == MS_FOR || g.kind == MS_K_WISE_NS ||
        g.kind == MS_GROUP_SIMPLE)
        return count / g.inner_bs + 2;
    return 0;
}

// Capacities of a blockwise_individual_f column: its index entries,
// reference slots and view-table bytes.
struct col_caps {
    uint64_t index_cap, refs_cap, tab_bytes;
};

// Every live column handle: ms_col_free refuses a handle it did not hand
// out (or already freed) instead of freeing it twice.
static std::mutex g_cols_mu;
static std::unordered_set<const ms_col *> g_cols;

// One allocation: [meta 256 B][index][refs][view table][image]
// A new column; its meta is cleared on the device unless the caller's
// first kernel writes all of it (clear_meta false).
static int new_col(const ms_format_desc &f, uint64_t count_bound, uint64_t size_alloc, hipStream_t s, ms_col **out,
                   const col_caps *caps = nullptr, bool clear_meta = true)
{
    ms_col *c = new (std::nothrow) ms_col();
    if (!c) return ms_fail(MS_ERR_NOMEM, "host allocation");
    c->f = f;
    c->g = ms_make_geom(f);
    c->count_bound = count_bound;
    c->stream = s;
    c->index_cap = caps ? caps->index_cap : idx_entries(c->g, count_bound);
    c->refs_cap = caps ? caps->refs_cap
                : f.kind == MS_FOR ? count_bound / f.block + 2
                : f.kind == MS_DELTA ? (count_bound / MS_DELTA_TILE + 2) * 8
                : f.kind == MS_GROUP_SIMPLE ? count_bound / f.step + 2 : 0;
    c->size_alloc = size_alloc;
    const uint64_t idx_bytes = (c->index_cap * 8 + 255) / 256 * 256;
    const uint64_t refs_bytes = (c->refs_cap * 8 + 255) / 256 * 256;
    const uint64_t tab_bytes = caps ? (caps->tab_bytes + 255) / 256 * 256 : 0;
    const uint64_t data_bytes = (size_alloc + 1024 + 255) / 256 * 256;
    uint8_t *base = (uint8_t *)dev_alloc(256 + idx_bytes + refs_bytes + tab_bytes + data_bytes, s);
    if (!base) {
        delete c;
        return ms_fail(MS_ERR_NOMEM, "device allocation of %llu bytes", (unsigned long long)(size_alloc + 1024));
    }
    c->arena = g_rec && g_rec->recording && g_rec->mode == 1 ? g_rec : nullptr;
    c->meta = (ms_dev_meta *)base;
    c->index = c->index_cap ? (uint64_t *)(base + 256) : nullptr;
    c->refs = c->refs_cap ? (uint64_t *)(base + 256 + idx_bytes) : nullptr;
    c->vmeta = tab_bytes ? (ms_dev_meta *)(base + 256 + idx_bytes + refs_bytes) : nullptr;
    c->data = base + 256 + idx_bytes + refs_bytes + tab_bytes;
    int rc = clear_meta ? launch_fill_u64((uint64_t *)base, 256 / 8, 0, s) : MS_OK;
    if (rc) {
        dev_free(base, s);
        delete c;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(g_cols_mu);
        g_cols.insert(c);
    }
    *out = c;
    return MS_OK;
}

// Bytes a write-iterator with m_CountBuffer = K (0: a column-level morph,
// one morph_batch call) writes for at most m values.  Every K-value call
// starts fresh dynamic_vbp meta pages and for_f reference pages
// (write_iterator.h:78-185), so the calls are bounded one by one: at most
// m / K full calls and one done() call of fewer than K values, each at most
// F::get_size_max_byte(K) (monotone in its count); then the rest gap, its
// 64-byte alignment and a raw rest of fewer than m_BlockSize values.
static uint64_t wit_size_max(const ms_format_desc &f, uint64_t m, uint64_t K)
{
    if (f.kind == MS_UNCOMPR) return 8 * m;
    if (f.kind == MS_RLE) return 16 * m;
    const ms_geom g = ms_make_geom(f);
    if (K == 0) return size_max_any_len(f, m) + g.gap;
    return (m / K + 1) * size_max(f, K) + g.gap + 64 + 8 * g.bs;
}

static uint64_t out_capacity(const ms_format_desc &f, uint64_t count_bound, uint64_t K)
{
    // plus room for the raw appends of select/project behind the rest, and
    // for the zero padding of the rest to two blocks that a preparation
    // for random access writes
    const ms_geom g = ms_make_geom(f);
    return wit_size_max(f, count_bound, K) + 8 * 64 + 256 + 16 * g.bs;
}

static hipStream_t S(void *stream) { return (hipStream_t)stream; }

// Builds the block index of an image on the host (the producer-side kernels
// build it on the device).
static int host_index_walk(const ms_format_desc &f, const ms_geom &g, const uint8_t *img, const ms_col_meta &m,
                           std::vector<uint64_t> &idx, std::vector<uint64_t> &refs, uint64_t *walked)
{
    struct set_end { uint64_t *w; const uint64_t &p; ~set_end() { if (w) *w = p; } };
    const ms_parts pt = ms_parts_of(g, m.count, m.size_used, m.size_compr);
    const uint64_t end = m.size_compr;
    uint64_t pos = 0;
    const set_end report{walked, pos};   // the byte the walk ended at
    auto leaf_call = [&](uint64_t n) -> int {
        // one morph_batch call of the leaf layout over n values at pos
        if (g.leaf == MS_UNCOMPR) {
            if (pos + 8 * n > end) return ms_fail(MS_ERR_INVALID, "image too short");
            for (uint64_t i = 0; i < n; i += g.inner_bs) { idx.push_back((pos << 8) | 64); pos += 8 * g.inner_bs; }
            return MS_OK;
        }
        if (g.leaf == MS_STATIC_VBP) {
            for (uint64_t i = 0; i < n; i += g.inner_bs) {
                idx.push_back((pos << 8) | g.bw);
                pos += g.inner_bs * g.bw / 8;
            }
            return pos <= end ? MS_OK : ms_fail(MS_ERR_INVALID, "image too short");
        }
        uint64_t done = 0;
        while (done < n) {
            if (pos + g.leaf_page > end) return ms_fail(MS_ERR_INVALID, "dynamic_vbp page header past sizeCompr");
            const uint8_t *meta = img + pos;
            pos += g.leaf_page;
            for (uint64_t q = 0; q < g.leaf_page; q++) {
                const uint32_t bw = meta[q];
                if (bw == MS_VBP_BW_NOBLOCK) continue;
                if (bw < 1 || bw > 64) return ms_fail(MS_ERR_INVALID, "bit width %u in a page header", bw);
                idx.push_back((pos << 8) | bw);
                pos += g.inner_bs * bw / 8;
                done += g.inner_bs;
            }
            if (pos > end) return ms_fail(MS_ERR_INVALID, "dynamic_vbp block past sizeCompr");
        }
        return MS_OK;
    };
    switch (f.kind) {
    case MS_K_WISE_NS: {   // a byte offset per 32 pairs
        for (uint64_t i = 0; i < pt.n_compr / 2; i++) {
            if (pos >= end) return ms_fail(MS_ERR_INVALID, "k_wise_ns pair past sizeCompr");
            if (i % (MS_KWNS_PER_ENTRY / 2) == 0) idx.push_back(pos);
            const uint32_t d = img[pos];
            pos += 17 - (d & 7) - ((d >> 3) & 7);
        }
        return pos <= end ? MS_OK : ms_fail(MS_ERR_INVALID, "k_wise_ns data past sizeCompr");
    }
    case MS_GROUP_SIMPLE: {   // per block (data offset << 8) | code and its first value
        static const uint8_t num[15] = {64, 32, 21, 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1};
        const uint64_t E = f.step, a = 8 * E;
        uint64_t done = 0;
        while (done < pt.n_compr) {
            if (pos + 9 > end) return ms_fail(MS_ERR_INVALID, "group_simple header past sizeCompr");
            uint64_t cs;
            memcpy(&cs, img + pos, 8);
            const uint32_t R = img[pos + 8];
            if (R >= 64 || cs > pt.n_compr || pos + 9 + (cs + 1) / 2 > end)
                return ms_fail(MS_ERR_INVALID, "group_simple header");
            uint64_t w = (pos + 9 + (cs + 1) / 2 + a - 1) / a * a;
            for (uint64_t k = 0; k < cs; k++) {
                const uint32_t sl = (img[pos + 9 + k / 2] >> (4 * (k % 2))) & 15;
                if (sl > 14) return ms_fail(MS_ERR_INVALID, "group_simple selector %u", sl);
                idx.push_back((w << 8) | sl);
                refs.push_back(done);
                done += num[sl] * E;
                w += a;
            }
            if (R) {
                idx.push_back((w << 8) | (16 + R));
                refs.push_back(done);
                done += R * E;
                w += a;
            }
            if (!cs && !R) return ms_fail(MS_ERR_INVALID, "group_simple call without blocks");
            if (w > end || done > pt.n_compr) return ms_fail(MS_ERR_INVALID, "group_simple blocks past sizeCompr");
            pos = w;
        }
        return MS_OK;
    }
    case MS_DYNAMIC_VBP: return leaf_call(pt.n_compr);
    case MS_DELTA:
        for (uint64_t b = 0; b < pt.n_compr / f.block; b++) MS_TRY(leaf_call(f.block));
        return MS_OK;
    case MS_FOR: {
        uint64_t done = 0;
        while (done < pt.n_compr) {
            if (pos + 8ull * f.page > end) return ms_fail(MS_ERR_INVALID, "for_f page past sizeCompr");
            const uint64_t *r = (const uint64_t *)(img + pos);
            pos += 8ull * f.page;
            for (uint64_t q = 0; q < f.page && done < pt.n_compr; q++) {
                if (r[q] == ~0ull) break;
                refs.push_back(r[q]);
                MS_TRY(leaf_call(f.block));
                done += f.block;
            }
        }
        return MS_OK;
    }
    default: return MS_OK;
    }
}

static int host_index(const ms_format_desc &f, const ms_geom &g, const uint8_t *img, const ms_col_meta &m,
                      std::vector<uint64_t> &idx, std::vector<uint64_t> &refs, uint64_t *walked = nullptr)
{
    return host_index_walk(f, g, img, m, idx, refs, walked);
}

// host_index_walk over one blockwise_individual_f view: `count` values in
// at most `avail` bytes; *walked = the view's bytes.
static int view_walk(const ms_format_desc &d, const uint8_t *img, uint64_t count, uint64_t avail,
                     std::vector<uint64_t> &idx, std::vector<uint64_t> &refs, uint64_t *walked)
{
    const ms_col_meta m = {count, avail, avail, avail};
    return host_index_walk(d, ms_make_geom(d), img, m, idx, refs, walked);
}

static int check_meta(const ms_format_desc &f, const ms_col_meta &m)
{
    const ms_geom g = ms_make_geom(f);
    if (m.size_compr > m.size_used) return ms_fail(MS_ERR_INVALID, "sizeCompr > sizeUsed");
    if (f.kind == MS_UNCOMPR) {
        if (m.size_used != 8 * m.count) return ms_fail(MS_ERR_INVALID, "uncompr_f: sizeUsed != 8*count");
        return MS_OK;
    }
    if (f.kind == MS_RLE) {
        if (m.size_used % 16 || m.size_used != m.size_compr) return ms_fail(MS_ERR_INVALID, "rle_f: bad sizes");
        return MS_OK;
    }
    if (m.size_used != m.size_compr) {
        const uint64_t r = ms_rest_offset(g, m.size_compr);
        if (m.size_used < r || (m.size_used - r) % 8) return ms_fail(MS_ERR_INVALID, "bad rest placement");
    }
    const ms_parts pt = ms_parts_of(g, m.count, m.size_used, m.size_compr);
    if (pt.n_rest > m.count) return ms_fail(MS_ERR_INVALID, "rest larger than count");
    if (pt.n_compr % g.bs) return ms_fail(MS_ERR_INVALID, "compressed part not a multiple of the block size");
    if (f.kind == MS_STATIC_VBP && m.size_compr != pt.n_compr * f.bw / 8)
        return ms_fail(MS_ERR_INVALID, "static_vbp: sizeCompr does not match count");
    if (f.kind == MS_STATIC_VBP_PAD && m.size_compr != pt.n_compr / g.bs * f.step * 8)
        return ms_fail(MS_ERR_INVALID, "static_vbp<vbp_padding_l>: sizeCompr does not match count");
    return MS_OK;
}

extern "C" int ms_col_upload(ms_col **out, const ms_format_desc *f, const void *image, const ms_col_meta *meta,
                             void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !meta || (!image && meta->size_used)) return ms_fail(MS_ERR_INVALID, "null argument");
    if (!valid_format(f)) return ms_fail(MS_ERR_INVALID, "invalid format");
    MS_TRY(check_meta(*f, *meta));
    const ms_geom g = ms_make_geom(*f);
    if (f->kind == MS_BLOCKWISE_INDIVIDUAL) {
        bwi_host hb;
        MS_TRY(bwi_walk_image(*f, (const uint8_t *)image, meta->count, meta->size_compr, hb, view_walk));
        if (hb.h.C != meta->size_compr)
            return ms_fail(MS_ERR_INVALID, "blockwise_individual_f: sizeCompr %llu, the views end at %llu",
                           (unsigned long long)meta->size_compr, (unsigned long long)hb.h.C);
        const uint64_t alloc = std::max(meta->size_used, ms_rest_offset(g, meta->size_compr) + 8 * g.bs) + 64;
        const col_caps caps = {hb.idx.size() + 1, bwi_refs_entries(hb.h) + 8, bwi_tab_bytes(hb.h.nb)};
        ms_col *c;
        MS_TRY(new_col(*f, meta->count, alloc, S(stream), &c, &caps));
        ms_dev_meta dm = {};
        dm.count = meta->count;
        dm.size_used = meta->size_used;
        dm.size_compr = meta->size_compr;
        temps sc(S(stream));
        int rc = MS_OK;
        if (meta->size_used && hipMemcpyAsync(c->data, image, meta->size_used, hipMemcpyHostToDevice, S(stream)))
            rc = ms_fail(MS_ERR_HIP, "upload");
        if (!rc && !hb.idx.empty() &&
            hipMemcpyAsync(c->index, hb.idx.data(), hb.idx.size() * 8, hipMemcpyHostToDevice, S(stream)))
            rc = ms_fail(MS_ERR_HIP, "upload");
        if (!rc && !hb.refs.empty() &&
            hipMemcpyAsync(c->refs, hb.refs.data(), hb.refs.size() * 8, hipMemcpyHostToDevice, S(stream)))
            rc = ms_fail(MS_ERR_HIP, "upload");
        if (!rc && hipMemcpyAsync(c->meta, &dm, sizeof dm, hipMemcpyHostToDevice, S(stream)))
            rc = ms_fail(MS_ERR_HIP, "upload");
        if (!rc) rc = bwi_finish_upload(c, hb, sc);
        if (!rc && hipStreamSynchronize(S(stream))) rc = ms_fail(MS_ERR_HIP, "upload");
        if (rc) { ms_col_free(c); return rc; }
        *out = c;
        return MS_OK;
    }
    std::vector<uint64_t> idx, refs;
    MS_TRY(host_index(*f, g, (const uint8_t *)image, *meta, idx, refs));
    ms_col *c;
    uint64_t alloc = meta->size_used;
    if (f->kind != MS_UNCOMPR && f->kind != MS_RLE)   // room for a preparation for random access too
        alloc = std::max(alloc, size_max_any_len(*f, meta->count)) + 16 * g.bs + 64;
    MS_TRY(new_col(*f, meta->count, alloc, S(stream), &c));
    if (idx.size() > c->index_cap || refs.size() > c->refs_cap) {
        ms_col_free(c);
        return ms_fail(MS_ERR_CAPACITY, "index larger than expected");
    }
    ms_dev_meta dm = {};
    dm.count = meta->count;
    dm.size_used = meta->size_used;
    dm.size_compr = meta->size_compr;
    if (f->kind == MS_GROUP_SIMPLE) dm.aux[2] = idx.size();   // its blocks
    if (meta->size_used) MS_HIP(hipMemcpyAsync(c->data, image, meta->size_used, hipMemcpyHostToDevice, S(stream)));
    if (!idx.empty()) MS_HIP(hipMemcpyAsync(c->index, idx.data(), idx.size() * 8, hipMemcpyHostToDevice, S(stream)));
    if (!refs.empty()) MS_HIP(hipMemcpyAsync(c->refs, refs.data(), refs.size() * 8, hipMemcpyHostToDevice, S(stream)));
    MS_HIP(hipMemcpyAsync(c->meta, &dm, sizeof dm, hipMemcpyHostToDevice, S(stream)));
    if (f->kind == MS_DELTA) {
        temps sc(S(stream));
        int rc = launch_delta_carries(c, sc);
        if (rc) { ms_col_free(c); return rc; }
    }
    MS_HIP(hipStreamSynchronize(S(stream)));
    *out = c;
    return MS_OK;
}

extern "C" int ms_col_upload_values(ms_col **out, const uint64_t *values, uint64_t count, void *stream)
{
    ms_format_desc f = {};
    f.kind = MS_UNCOMPR;
    ms_col_meta m = {count, 8 * count, 0, 8 * count};
    return ms_col_upload(out, &f, values, &m, stream);
}

extern "C" int ms_col_from_device(ms_col **out, const uint64_t *d_values, uint64_t count, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || (!d_values && count)) return ms_fail(MS_ERR_INVALID, "null argument");
    ms_format_desc f = {};
    f.kind = MS_UNCOMPR;
    ms_col *c;
    MS_TRY(new_col(f, count, 8 * count, S(stream), &c));
    if (count) MS_HIP(hipMemcpyAsync(c->data, d_values, 8 * count, hipMemcpyDeviceToDevice, S(stream)));
    ms_dev_meta dm = {};
    dm.count = count;
    dm.size_used = 8 * count;
    MS_HIP(hipMemcpyAsync(c->meta, &dm, sizeof dm, hipMemcpyHostToDevice, S(stream)));
    MS_HIP(hipStreamSynchronize(S(stream)));   // dm lives on this stack frame
    *out = c;
    return MS_OK;
}

extern "C" int ms_col_get_meta(const ms_col *c, ms_col_meta *meta, void *stream)
{
    if (!c || !meta) return ms_fail(MS_ERR_INVALID, "null argument");
    ms_dev_meta dm;
    MS_HIP(hipMemcpyAsync(&dm, c->meta, sizeof dm, hipMemcpyDeviceToHost, S(stream)));
    MS_HIP(hipStreamSynchronize(S(stream)));
    meta->count = dm.count;
    meta->size_used = dm.size_used;
    meta->size_compr = dm.size_compr;
    meta->size_alloc = c->size_alloc;
    return MS_OK;
}

extern "C" int ms_col_check(const ms_col *c, void *stream)
{
    if (!c) return ms_fail(MS_ERR_INVALID, "null column");
    ms_dev_meta dm;
    MS_HIP(hipMemcpyAsync(&dm, c->meta, sizeof dm, hipMemcpyDeviceToHost, S(stream)));
    MS_HIP(hipStreamSynchronize(S(stream)));
    if (dm.status) return ms_fail((int)dm.status, "a kernel producing this column reported status %llu",
                                  (unsigned long long)dm.status);
    return MS_OK;
}

extern "C" int ms_col_download(const ms_col *c, void *image, uint64_t capacity, ms_col_meta *meta, void *stream)
{
    ms_col_meta m;
    MS_TRY(ms_col_get_meta(c, &m, stream));
    if (meta) *meta = m;
    if (!image) return MS_OK;
    if (capacity < m.size_used) return ms_fail(MS_ERR_CAPACITY, "download buffer of %llu < sizeUsed %llu",
                                               (unsigned long long)capacity, (unsigned long long)m.size_used);
    if (m.size_used == 0) return MS_OK;
    const ms_parts pt = ms_parts_of(c->g, m.count, m.size_used, m.size_compr);
    if (c->f.kind == MS_UNCOMPR || c->f.kind == MS_RLE || m.size_used == m.size_compr) {
        MS_HIP(hipMemcpyAsync(image, c->data, m.size_used, hipMemcpyDeviceToHost, S(stream)));
    } else {
        // the gap between the compressed part and the rest reads as zeros,
        // but for the compressed rest a preparation for random access put
        // at its start
        uint8_t *h = (uint8_t *)image;
        const uint64_t head = m.size_compr + c->prepared_bytes;
        memset(h + head, 0, pt.rest_off - head);
        if (head) MS_HIP(hipMemcpyAsync(h, c->data, head, hipMemcpyDeviceToHost, S(stream)));
        MS_HIP(hipMemcpyAsync(h + pt.rest_off, c->data + pt.rest_off, m.size_used - pt.rest_off,
                              hipMemcpyDeviceToHost, S(stream)));
    }
    MS_HIP(hipStreamSynchronize(S(stream)));
    return MS_OK;
}

extern "C" int ms_col_format(const ms_col *c, ms_format_desc *f)
{
    if (!c || !f) return ms_fail(MS_ERR_INVALID, "null argument");
    *f = c->f;
    return MS_OK;
}

extern "C" const void *ms_col_device_data(const ms_col *c) { return c ? c->data : nullptr; }
extern "C" uint64_t ms_col_count_bound(const ms_col *c) { return c ? c->count_bound : 0; }

extern "C" int ms_col_free(ms_col *c)
{
    if (!c) return MS_OK;
    {
        std::lock_guard<std::mutex> lk(g_cols_mu);
        if (!g_cols.erase(c)) return ms_fail(MS_ERR_INVALID, "ms_col_free: not a live column handle");
    }
    if (c->arena) {
        // arena columns die with their plan
    } else if (g_rec && g_rec->recording && g_rec->mode == 1) {
        g_rec->deferred.push_back((void *)c->meta);   // a free now would be captured into the graph
    } else {
        (void)hipFreeAsync((void *)c->meta, c->stream);
    }
    delete c->bwi;
    delete c;
    return MS_OK;
}

static int emit_new(const ms_format_desc &f, const uint64_t *vals, const uint64_t *d_m, uint64_t m_bound,
                    uint64_t chunk, uint64_t extra_values, temps &sc, ms_col **out, uint64_t **d_append);

// column<F>::prepare_for_random_access (column.h:241-268): the rest,
// zero-padded to a multiple of the block size, compressed by one fresh
// morph_batch call to the end of the compressed part (inside the gap).
extern "C" int ms_col_prepare_for_random_access(ms_col *c, int *acted, void *stream)
{
    MS_TRY(ms_need_init());
    if (!c) return ms_fail(MS_ERR_INVALID, "null column");
    if (acted) *acted = 0;
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "prepare_for_random_access synchronises: not in a plan");
    if (c->f.kind == MS_RLE) return ms_fail(MS_ERR_INVALID, "rle_f has no block size");
    if (c->prepared || c->f.kind == MS_UNCOMPR) return MS_OK;
    ms_col_meta m;
    MS_TRY(ms_col_get_meta(c, &m, stream));
    const ms_parts pt = ms_parts_of(c->g, m.count, m.size_used, m.size_compr);
    if (!pt.n_rest) return MS_OK;
    const uint64_t bs = c->g.bs;
    const uint64_t ru = (pt.n_rest + bs - 1) / bs * bs;
    if (pt.rest_off + 8 * ru > c->size_alloc)
        return ms_fail(MS_ERR_CAPACITY, "no room to pad the rest of %llu values to %llu",
                       (unsigned long long)pt.n_rest, (unsigned long long)ru);
    temps sc(S(stream));
    uint64_t *rest = (uint64_t *)(c->data + pt.rest_off);
    MS_TRY(launch_fill_u64(rest + pt.n_rest, ru - pt.n_rest, 0, S(stream)));
    if (c->f.kind == MS_BLOCKWISE_INDIVIDUAL) {
        // one block, formatted by the column's selector, its code byte at
        // sizeCompr and its padding aligned to the image (create_aligned_ptr)
        bwi_work w;
        MS_TRY(bwi_plan(c->f, rest, ru, nullptr, m.size_compr, w, sc));
        uint64_t *ti = (uint64_t *)sc.get((bwi_index_entries(w.h) + 1) * 8 + 64);
        uint64_t *tr = (uint64_t *)sc.get((bwi_refs_entries(w.h) + 8) * 8 + 64);
        if (!ti || !tr) return ms_fail(MS_ERR_NOMEM, "prepare scratch");
        if (w.h.C > pt.rest_off) return ms_fail(MS_ERR_CAPACITY, "compressed rest does not fit the gap");
        MS_TRY(bwi_write_image(w, c->data, pt.rest_off, &c->meta->status, ti, tr, sc));
        MS_HIP(hipStreamSynchronize(S(stream)));
        c->prepared = true;
        c->prepared_bytes = w.h.C - m.size_compr;
        if (acted) *acted = 1;
        return MS_OK;
    }
    uint64_t *d_m = (uint64_t *)sc.get(64);
    if (!d_m) return ms_fail(MS_ERR_NOMEM, "prepare scratch");
    MS_TRY(launch_fill_u64(d_m, 1, ru, S(stream)));
    ms_col *t;
    uint64_t *app;
    MS_TRY(emit_new(c->f, rest, d_m, ru, 0, 0, sc, &t, &app));
    ms_col_meta tm;
    int rc = ms_col_get_meta(t, &tm, stream);
    if (!rc && (tm.size_compr != tm.size_used || m.size_compr + tm.size_compr > pt.rest_off))
        rc = ms_fail(MS_ERR_CAPACITY, "compressed rest does not fit the gap");
    if (!rc && tm.size_compr &&
        hipMemcpyAsync(c->data + m.size_compr, t->data, tm.size_compr, hipMemcpyDeviceToDevice, S(stream)) !=
            hipSuccess)
        rc = ms_fail(MS_ERR_HIP, "prepare copy");
    if (!rc && hipStreamSynchronize(S(stream)) != hipSuccess) rc = ms_fail(MS_ERR_HIP, "prepare sync");
    ms_col_free(t);
    if (rc) return rc;
    c->prepared = true;
    c->prepared_bytes = tm.size_compr;
    if (acted) *acted = 1;
    return MS_OK;
}

// ------------------------------------------------------------------ binary_io
// binary_io<F>::store / load (binary_io.h:43-87): the file holds the count
// and sizeUsed as two little-endian uint64, then the sizeUsed bytes of the
// image.  The file has no sizeCompr: the reference's loaded column keeps the
// constructor's 0 (right for uncompr_f only); here the loader derives it
// from the image as a column-level morph of `count` values lays it out
// (floor(count / block) values compressed, the rest raw behind the gap;
// rle_f: all of it) and refuses an image that does not fit that layout.
extern "C" int ms_col_store(const ms_col *c, const char *path, void *stream)
{
    if (!c || !path) return ms_fail(MS_ERR_INVALID, "null argument");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "binary_io synchronises: not in a plan");
    ms_col_meta m;
    MS_TRY(ms_col_get_meta(c, &m, stream));
    std::vector<uint8_t> img(m.size_used + 1);
    MS_TRY(ms_col_download(c, img.data(), img.size(), nullptr, stream));
    FILE *fp = fopen(path, "wb");
    if (!fp) return ms_fail(MS_ERR_INVALID, "could open the file for writing: %s", path);
    const uint64_t hdr[2] = {m.count, m.size_used};
    const bool ok = fwrite(hdr, 8, 2, fp) == 2 && fwrite(img.data(), 1, m.size_used, fp) == m.size_used;
    const bool closed = fclose(fp) == 0;
    if (!ok || !closed) return ms_fail(MS_ERR_INVALID, "could write the column meta data and data: %s", path);
    return MS_OK;
}

extern "C" int ms_col_load(ms_col **out, const ms_format_desc *f, const char *path, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !path) return ms_fail(MS_ERR_INVALID, "null argument");
    if (!valid_format(f)) return ms_fail(MS_ERR_INVALID, "invalid format");
    FILE *fp = fopen(path, "rb");
    if (!fp) return ms_fail(MS_ERR_INVALID, "could not open the file for reading: %s", path);
    uint64_t hdr[2];
    if (fread(hdr, 8, 2, fp) != 2) {
        fclose(fp);
        return ms_fail(MS_ERR_INVALID, "could not read the column meta data: %s", path);
    }
    std::vector<uint8_t> img;
    try {
        img.resize(hdr[1] + 1);
    } catch (...) {
        fclose(fp);
        return ms_fail(MS_ERR_NOMEM, "%llu bytes for the column data", (unsigned long long)hdr[1]);
    }
    const bool ok = fread(img.data(), 1, hdr[1], fp) == hdr[1];
    fclose(fp);
    if (!ok) return ms_fail(MS_ERR_INVALID, "could not read the column data: %s", path);
    ms_col_meta m = {hdr[0], hdr[1], 0, hdr[1]};
    const ms_geom g = ms_make_geom(*f);
    if (f->kind == MS_RLE) {
        m.size_compr = m.size_used;
    } else if (f->kind != MS_UNCOMPR) {
        const uint64_t nc = m.count / g.bs * g.bs, rest = m.count - nc;
        uint64_t end = 0;
        std::vector<uint64_t> idx, refs;
        if (f->kind == MS_STATIC_VBP) {
            end = nc * f->bw / 8;
        } else if (f->kind == MS_BLOCKWISE_INDIVIDUAL) {
            bwi_host hb;
            MS_TRY(bwi_walk_image(*f, img.data(), nc, m.size_used, hb, view_walk));
            end = hb.h.C;
        } else {
            const ms_col_meta walk = {nc, m.size_used, m.size_used, m.size_used};
            MS_TRY(host_index(*f, g, img.data(), walk, idx, refs, &end));
        }
        const uint64_t want = rest ? ms_rest_offset(g, end) + 8 * rest : end;
        if (want != m.size_used)
            return ms_fail(MS_ERR_INVALID, "the image does not have a column-level morph's layout of %llu values",
                           (unsigned long long)m.count);
        m.size_compr = end;
    }
    return ms_col_upload(out, f, img.data(), &m, stream);
}

// ------------------------------------------------------------------ plans
extern "C" int ms_plan_create(ms_plan **out, uint64_t arena_bytes)
{
    MS_TRY(ms_need_init());
    if (!out) return ms_fail(MS_ERR_INVALID, "null argument");
    ms_plan *p = new (std::nothrow) ms_plan();
    if (!p) return ms_fail(MS_ERR_NOMEM, "host allocation");
    p->arena_bytes = arena_bytes;
    p->mode = arena_bytes ? 1 : 0;
    if (arena_bytes && hipMalloc((void **)&p->arena, arena_bytes) != hipSuccess) {
        delete p;
        return ms_fail(MS_ERR_NOMEM, "plan arena of %llu bytes", (unsigned long long)arena_bytes);
    }
    *out = p;
    return MS_OK;
}

extern "C" int ms_plan_begin(ms_plan *p, void *stream)
{
    if (!p || !stream) return ms_fail(MS_ERR_INVALID, "a plan records on a created (non-default) stream");
    if (g_rec) return ms_fail(MS_ERR_INVALID, "another plan is recording on this thread");
    if (p->mode == 0 && p->measured) {
        // the measuring pass is done: allocate the arena and capture
        p->arena_bytes = p->measured + (1u << 20);
        if (hipMalloc((void **)&p->arena, p->arena_bytes) != hipSuccess)
            return ms_fail(MS_ERR_NOMEM, "plan arena of %llu bytes", (unsigned long long)p->arena_bytes);
        p->mode = 1;
    }
    if (p->exec) return ms_fail(MS_ERR_INVALID, "plan already captured");
    p->stream = S(stream);
    p->used = 0;
    p->recording = true;
    g_rec = p;
    if (p->mode == 1) {
        MS_HIP(hipDeviceSynchronize());
        hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            g_rec = nullptr;
            p->recording = false;
            return ms_fail(MS_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
        }
    }
    return MS_OK;
}

extern "C" int ms_plan_end(ms_plan *p, void *stream)
{
    if (!p || g_rec != p || S(stream) != p->stream) return ms_fail(MS_ERR_INVALID, "plan is not recording here");
    g_rec = nullptr;
    p->recording = false;
    if (p->mode == 0) {
        MS_HIP(hipStreamSynchronize(p->stream));
        return MS_OK;
    }
    hipError_t e = hipStreamEndCapture(p->stream, &p->graph);
    for (void *q : p->deferred) (void)hipFreeAsync(q, p->stream);
    p->deferred.clear();
    if (e != hipSuccess) return ms_fail(MS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    MS_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return MS_OK;
}

extern "C" int ms_plan_launch(ms_plan *p, void *stream)
{
    if (!p || !p->exec) return ms_fail(MS_ERR_INVALID, "plan has no captured graph");
    MS_HIP(hipGraphLaunch(p->exec, S(stream)));
    return MS_OK;
}

extern "C" uint64_t ms_plan_arena_bytes(const ms_plan *p) { return p ? (p->mode ? p->arena_bytes : p->measured) : 0; }

extern "C" int ms_plan_free(ms_plan *p)
{
    if (!p) return MS_OK;
    if (g_rec == p) g_rec = nullptr;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (auto &h : p->holds) h.first(h.second);
    if (p->arena) {
        (void)hipDeviceSynchronize();
        (void)hipFree(p->arena);
    }
    delete p;
    return MS_OK;
}

// ------------------------------------------------------------------ helpers
// Uncompressed values of `in` (its own image for uncompr_f, else a
// decompressed temporary) and a device pointer to their count.
static int values_of(const ms_col *in, temps &sc, const uint64_t **vals, const uint64_t **d_n)
{
    *d_n = &in->meta->count;
    if (in->f.kind == MS_UNCOMPR) {
        *vals = (const uint64_t *)in->data;
        return MS_OK;
    }
    uint64_t *tmp = (uint64_t *)sc.get(in->count_bound * 8 + 64);
    if (!tmp) return ms_fail(MS_ERR_NOMEM, "decompression scratch");
    MS_TRY(launch_decompress(in, tmp, sc.s));
    *vals = tmp;
    return MS_OK;
}

static uint64_t wit_chunk(const ms_format_desc &f)
{
    const uint64_t bs = ms_make_geom(f).bs;
    return (bs + 2047) / 2048 * 2048;            // write_iterator.h:78-80
}

// The write-iterator flushes m_CountBuffer values per morph_batch call; the
// formats it supports here have blocks that tile that buffer.
static bool wit_ok(const ms_format_desc *f)
{
    if (f->kind == MS_UNCOMPR) return true;
    if (f->kind == MS_RLE || xfmt(f->kind) || f->kind == MS_BLOCKWISE_INDIVIDUAL) return false;
    const uint64_t bs = ms_make_geom(*f).bs;
    return wit_chunk(*f) % bs == 0;
}

// uncompr_f write-iterator (uncompr.h:125-221): a copy; done() returns
// (0, end, end).
__global__ void k_emit_uncompr(const uint64_t *v, const uint64_t *d_m, ms_dev_meta *meta, uint64_t *out)
{
    const uint64_t m = *d_m;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = v[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        meta->count = m;
        meta->size_used = 8 * m;
        meta->size_compr = 0;
        meta->aux[0] = 8 * m;
    }
}

static uint32_t grid_of(uint64_t n)
{
    uint64_t g = (n + 1023) / 1024;
    return (uint32_t)(g < 1 ? 1 : g > 8192 ? 8192 : g);
}

// Emits values[0 .. *d_m) of a write-iterator into a new column of format f
// (not rle_f); uncompr_f is a plain copy.
static int emit_new(const ms_format_desc &f, const uint64_t *vals, const uint64_t *d_m, uint64_t m_bound,
                    uint64_t chunk, uint64_t extra_values, temps &sc, ms_col **out, uint64_t **d_append)
{
    ms_col *c;
    MS_TRY(new_col(f, m_bound + extra_values, out_capacity(f, m_bound + extra_values, chunk), sc.s, &c));
    *d_append = &c->meta->aux[0];
    int rc;
    if (f.kind == MS_UNCOMPR) {
        hipLaunchKernelGGL(k_emit_uncompr, dim3(grid_of(m_bound)), dim3(256), 0, sc.s, vals, d_m, c->meta,
                           (uint64_t *)c->data);
        rc = hipGetLastError() == hipSuccess ? MS_OK : ms_fail(MS_ERR_HIP, "emit uncompr");
    } else if (xfmt(f.kind)) {
        rc = chunk == 0 ? launch_xfmt_compress(c, vals, d_m, m_bound, 0, sc)
                        : ms_fail(MS_ERR_INVALID, "no write-iterator for format kind %u", f.kind);
    } else {
        rc = launch_emit(c, vals, d_m, m_bound, chunk, nullptr, sc);
    }
    if (rc) {
        ms_col_free(c);
        return rc;
    }
    *out = c;
    return MS_OK;
}

// Device bookkeeping for the scalar remainder of select/project
// (select_compr.h:209-250, project_compr.h:183-252): the last r = rest mod
// lanes values of the input are processed scalar and their outputs appended
// raw behind the write-iterator's output.
__global__ void k_split_select(const ms_dev_meta *in, ms_geom g, uint32_t lanes, const uint64_t *pos,
                               const uint64_t *d_cnt, uint64_t *d_na, uint64_t *d_flag)
{
    const ms_parts pt = ms_parts_of(g, in->count, in->size_used, in->size_compr);
    const bool has_rest = g.kind == MS_UNCOMPR ? in->count != 0 : in->size_compr != in->size_used;
    const uint64_t r = has_rest ? pt.n_rest % lanes : 0;
    const uint64_t thr = pt.n - r;
    const uint64_t cnt = *d_cnt;
    uint64_t lo = 0, hi = cnt;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (pos[mid] < thr) lo = mid + 1; else hi = mid;
    }
    *d_na = lo;
    *d_flag = r != 0;
}

__global__ void k_split_project(const ms_dev_meta *pos, ms_geom g, uint32_t lanes, uint64_t *d_na, uint64_t *d_flag)
{
    const ms_parts pt = ms_parts_of(g, pos->count, pos->size_used, pos->size_compr);
    const bool has_rest = g.kind == MS_UNCOMPR ? pos->count != 0 : pos->size_compr != pos->size_used;
    const uint64_t r = has_rest ? pt.n_rest % lanes : 0;
    *d_na = pt.n - r;
    *d_flag = r != 0;
}

__global__ void k_append_raw(ms_dev_meta *m, uint8_t *data, uint64_t cap, const uint64_t *v, const uint64_t *d_na,
                             const uint64_t *d_total, const uint64_t *d_flag)
{
    const uint64_t na = *d_na, total = *d_total;
    const uint64_t app = m->aux[0];
    if (*d_flag && app + 8 * (total - na) > cap) {
        if (threadIdx.x == 0) m->status = MS_ERR_CAPACITY;
        return;
    }
    uint64_t *dst = (uint64_t *)(data + app);
    for (uint64_t i = na + threadIdx.x; i < total; i += blockDim.x) dst[i - na] = v[i];
    if (threadIdx.x == 0 && *d_flag) {
        m->count = total;
        m->size_used = app + 8 * (total - na);
    }
}

static int finish_split(ms_col *c, const uint64_t *v, const uint64_t *d_na, const uint64_t *d_total,
                        const uint64_t *d_flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_append_raw, dim3(1), dim3(64), 0, s, c->meta, c->data, c->size_alloc, v, d_na, d_total,
                       d_flag);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

static ms_format_desc uncompr_desc()
{
    ms_format_desc f = {};
    f.kind = MS_UNCOMPR;
    return f;
}

static int host_count(const ms_col *c, void *stream, uint64_t *n)
{
    ms_col_meta m;
    MS_TRY(ms_col_get_meta(c, &m, stream));
    *n = m.count;
    return MS_OK;
}

// ------------------------------------------------------------------ morph
// A blockwise_individual_f column of the n values at d_v, the block
// formats given by codes (host, one per block) or chosen by f's selector.
static int bwi_new(const ms_format_desc &f, const uint64_t *d_v, uint64_t n, const uint8_t *codes, temps &sc,
                   ms_col **out)
{
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "blockwise_individual_f compression synchronises: not in a plan");
    bwi_work w;
    MS_TRY(bwi_plan(f, d_v, n, codes, 0, w, sc));
    const ms_geom g = ms_make_geom(f);
    const uint64_t rest = n - w.h.nb * w.h.B;
    const uint64_t R = ms_align64(w.h.C + g.gap);
    // room for the rest zero-padded to a block (a preparation for random access)
    const uint64_t alloc = std::max(rest ? R + 8 * rest : w.h.C, R + 8 * g.bs) + 64;
    const col_caps caps = {bwi_index_entries(w.h) + 1, bwi_refs_entries(w.h) + 8, bwi_tab_bytes(w.h.nb)};
    ms_col *c;
    MS_TRY(new_col(f, n, alloc, sc.s, &c, &caps));
    int rc = bwi_pack(w, c, sc);
    if (rc) {
        ms_col_free(c);
        return rc;
    }
    *out = c;
    return MS_OK;
}

extern "C" int ms_morph(ms_col **out, const ms_format_desc *dst, const ms_col *in, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !in) return ms_fail(MS_ERR_INVALID, "null argument");
    if (!valid_format(dst)) return ms_fail(MS_ERR_INVALID, "invalid destination format");
    temps sc(S(stream));
    const uint64_t nb = in->count_bound;
    if (memcmp(dst, &in->f, sizeof *dst) == 0 || (dst->kind == MS_UNCOMPR && in->f.kind == MS_UNCOMPR)) {
        // same format: a copy of the image, meta, index and view table
        ms_col *c;
        const col_caps caps = {in->index_cap, in->refs_cap,
                               in->vmeta ? (uint64_t)(in->data - (uint8_t *)in->vmeta) : 0};
        MS_TRY(new_col(in->f, nb, in->size_alloc, S(stream), &c, in->bwi ? &caps : nullptr));
        const uint64_t bytes = (uint64_t)(c->data - (uint8_t *)c->meta) + in->size_alloc;
        MS_HIP(hipMemcpyAsync(c->meta, in->meta, bytes, hipMemcpyDeviceToDevice, S(stream)));
        c->prepared = in->prepared;
        c->prepared_bytes = in->prepared_bytes;
        if (in->bwi) {
            c->bwi = new (std::nothrow) ms_bwi(*in->bwi);
            if (!c->bwi) {
                ms_col_free(c);
                return ms_fail(MS_ERR_NOMEM, "host allocation");
            }
        }
        *out = c;
        return MS_OK;
    }
    ms_col *c;
    if (dst->kind == MS_UNCOMPR) {
        // decompression straight into the new column's buffer
        MS_TRY(new_col(*dst, nb, 8 * nb, S(stream), &c, nullptr, false));
        int rc = launch_decompress(in, (uint64_t *)c->data, S(stream), c->meta);   // writes the whole meta
        if (rc) { ms_col_free(c); return rc; }
        *out = c;
        return MS_OK;
    }
    const uint64_t *vals, *d_n;
    MS_TRY(values_of(in, sc, &vals, &d_n));
    if (dst->kind == MS_BLOCKWISE_INDIVIDUAL) {
        uint64_t n;
        MS_TRY(host_count(in, stream, &n));
        return bwi_new(*dst, vals, n, nullptr, sc, out);
    }
    if (dst->kind == MS_RLE) {
        MS_TRY(new_col(*dst, nb, 16 * nb + 64, S(stream), &c));
        int rc = launch_rle_compress(c, vals, d_n, nb, sc);
        if (rc) { ms_col_free(c); return rc; }
        *out = c;
        return MS_OK;
    }
    MS_TRY(new_col(*dst, nb, out_capacity(*dst, nb, 0), S(stream), &c));
    int rc = xfmt(dst->kind) ? launch_xfmt_compress(c, vals, d_n, nb, 0, sc) : launch_emit(c, vals, d_n, nb, 0, nullptr, sc);
    if (rc) { ms_col_free(c); return rc; }
    *out = c;
    return MS_OK;
}

extern "C" int ms_morph_blockwise(ms_col **out, const ms_format_desc *dst, const ms_col *in, const uint8_t *codes,
                                  uint64_t ncodes, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !in || (!codes && ncodes)) return ms_fail(MS_ERR_INVALID, "null argument");
    if (!valid_format(dst) || dst->kind != MS_BLOCKWISE_INDIVIDUAL)
        return ms_fail(MS_ERR_INVALID, "destination is not a blockwise_individual_f");
    uint64_t n;
    MS_TRY(host_count(in, stream, &n));
    if (ncodes != n / dst->block)
        return ms_fail(MS_ERR_INVALID, "%llu codes for %llu blocks", (unsigned long long)ncodes,
                       (unsigned long long)(n / dst->block));
    temps sc(S(stream));
    const uint64_t *vals, *d_n;
    MS_TRY(values_of(in, sc, &vals, &d_n));
    return bwi_new(*dst, vals, n, codes, sc, out);
}

extern "C" int ms_col_block_codes(const ms_col *c, uint8_t *codes, uint64_t cap, uint64_t *n)
{
    if (!c || !n) return ms_fail(MS_ERR_INVALID, "null argument");
    if (!c->bwi) return ms_fail(MS_ERR_INVALID, "not a blockwise_individual_f column");
    const ms_bwi &h = *c->bwi;
    *n = h.nb;
    if (codes)
        for (uint64_t b = 0; b < h.nb && b < cap; b++) codes[b] = (uint8_t)h.v[b].code;
    return MS_OK;
}

// ------------------------------------------------------------------ select
// Positions with p(x) through the selective write-iterator into out_f, the
// input rest's last (rest mod lanes) values processed "scalar" and their
// positions appended raw (select_compr.h:100-260, between_compr.h:107-270).
static int select_pred(ms_col **out, const ms_pred &p, const ms_format_desc *out_f, const ms_col *in, uint32_t lanes,
                       void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !in || lanes == 0 || lanes > 64) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    temps sc(S(stream));
    const uint64_t nb = in->count_bound;
    uint64_t *pos = (uint64_t *)sc.get(nb * 8 + 64);
    uint64_t *d = (uint64_t *)sc.get(64);
    if (!pos || !d) return ms_fail(MS_ERR_NOMEM, "select scratch");
    uint64_t *d_cnt = d, *d_na = d + 1, *d_flag = d + 2;
    if (select_flags_ok(in) && emit_flags_ok(*out_f)) {
        // positions go from the count pass's ballots straight into the
        // output format; pos only receives the few read by rank
        sel_flags sf;
        MS_TRY(launch_select_flags(in, p, sf, sc));
        const sel_split sp = {in->meta, in->g, lanes, d_cnt, d_na, d_flag};
        const uint64_t chunk = wit_chunk(*out_f);
        ms_col *c;
        // the emission's first kernel clears the meta
        MS_TRY(new_col(*out_f, nb + 64, out_capacity(*out_f, nb + 64, chunk), S(stream), &c, nullptr, false));
        const int rc = launch_emit_flags(c, sf, sp, pos, nb, chunk, sc);   // split, emission and raw append
        if (rc) { ms_col_free(c); return rc; }
        *out = c;
        return MS_OK;
    }
    MS_TRY(launch_select_positions(in, p, pos, d_cnt, sc));
    hipLaunchKernelGGL(k_split_select, dim3(1), dim3(1), 0, S(stream), in->meta, in->g, lanes, pos, d_cnt, d_na,
                       d_flag);
    MS_HIP(hipGetLastError());
    ms_col *c;
    uint64_t *app;
    MS_TRY(emit_new(*out_f, pos, d_na, nb, wit_chunk(*out_f), 64, sc, &c, &app));
    int rc = finish_split(c, pos, d_na, d_cnt, d_flag, S(stream));
    if (rc) { ms_col_free(c); return rc; }
    *out = c;
    return MS_OK;
}

extern "C" int ms_select(ms_col **out, uint32_t cmp, uint64_t val, const ms_format_desc *out_f, const ms_col *in,
                         uint32_t lanes, void *stream)
{
    if (cmp > MS_CMP_GE) return ms_fail(MS_ERR_INVALID, "comparison %u", cmp);
    return select_pred(out, make_pred(cmp, val), out_f, in, lanes, stream);
}

// between_t<t_op_lower, t_op_upper>: x passes when both comparisons hold;
// for the interval comparisons (all but NE) that is one interval.
extern "C" int ms_between(ms_col **out, uint32_t cmp_lower, uint64_t val_lower, uint32_t cmp_upper,
                          uint64_t val_upper, const ms_format_desc *out_f, const ms_col *in, uint32_t lanes,
                          void *stream)
{
    if (cmp_lower > MS_CMP_GE || cmp_upper > MS_CMP_GE || cmp_lower == MS_CMP_NE || cmp_upper == MS_CMP_NE)
        return ms_fail(MS_ERR_INVALID, "between takes the interval comparisons EQ, LT, LE, GT, GE");
    const ms_pred a = make_pred(cmp_lower, val_lower), b = make_pred(cmp_upper, val_upper);
    ms_pred p = {a.lo > b.lo ? a.lo : b.lo, a.hi < b.hi ? a.hi : b.hi, 0, 0};
    p.empty = a.empty || b.empty || p.lo > p.hi;
    if (p.empty) p.lo = 0, p.hi = ~0ull;
    return select_pred(out, p, out_f, in, lanes, stream);
}

extern "C" int ms_select_sum(uint64_t *d_result, uint32_t cmp, uint64_t val, const ms_col *in, void *stream)
{
    MS_TRY(ms_need_init());
    if (!d_result || !in || cmp > MS_CMP_GE) return ms_fail(MS_ERR_INVALID, "bad argument");
    return launch_select_sum(in, make_pred(cmp, val), d_result, false, S(stream));
}

extern "C" int ms_select_sum_noalloc(uint64_t *d_result, uint32_t cmp, uint64_t val, const ms_col *in, void *stream)
{
    MS_TRY(ms_need_init());
    if (!d_result || !in || cmp > MS_CMP_GE) return ms_fail(MS_ERR_INVALID, "bad argument");
    int dev = -1;
    MS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES || !g_scratch_ptr[dev])
        return ms_fail(MS_ERR_NODEVICE, "ms_init was not called for device %d", dev);
    const ss_buffers bufs = {g_scratch_ptr[dev], g_scratch_ptr[dev] + SS_MAX_PARTS, NOALLOC_TMP_VALUES};
    return launch_select_sum(in, make_pred(cmp, val), d_result, false, S(stream), &bufs);
}

// ------------------------------------------------------------------ project
extern "C" int ms_project(ms_col **out, const ms_format_desc *out_f, const ms_col *data, const ms_col *pos,
                          uint32_t lanes, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !data || !pos || lanes == 0 || lanes > 64) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    temps sc(S(stream));
    const uint64_t np = pos->count_bound;
    const uint64_t *pv, *d_np;
    MS_TRY(values_of(pos, sc, &pv, &d_np));
    // data without random access (delta_f, rle_f and the formats of the
    // extra codecs) is decompressed first
    ms_col tmpdata;
    const ms_col *dsrc = data;
    if (data->f.kind == MS_DELTA || data->f.kind == MS_RLE || xfmt(data->f.kind) ||
        data->f.kind == MS_BLOCKWISE_INDIVIDUAL) {
        const uint64_t *dv, *d_dn;
        MS_TRY(values_of(data, sc, &dv, &d_dn));
        ms_dev_meta *dm = (ms_dev_meta *)sc.get(sizeof(ms_dev_meta));
        if (!dm) return ms_fail(MS_ERR_NOMEM, "project scratch");
        MS_TRY(launch_set_meta(dm, d_dn, 8, 0, S(stream)));
        tmpdata = ms_col();
        tmpdata.f = uncompr_desc();
        tmpdata.g = ms_make_geom(tmpdata.f);
        tmpdata.data = (uint8_t *)dv;
        tmpdata.meta = dm;
        tmpdata.count_bound = data->count_bound;
        dsrc = &tmpdata;
    }
    uint64_t *v = (uint64_t *)sc.get(np * 8 + 64);
    uint64_t *d = (uint64_t *)sc.get(64);
    if (!v || !d) return ms_fail(MS_ERR_NOMEM, "project scratch");
    uint64_t *d_na = d, *d_flag = d + 1, *d_status = d + 2;
    MS_TRY(launch_fill_u64(d_status, 1, 0, S(stream)));
    MS_TRY(launch_gather(dsrc, pv, d_np, np, v, d_status, sc));
    hipLaunchKernelGGL(k_split_project, dim3(1), dim3(1), 0, S(stream), pos->meta, pos->g, lanes, d_na, d_flag);
    MS_HIP(hipGetLastError());
    ms_col *c;
    uint64_t *app;
    MS_TRY(emit_new(*out_f, v, d_na, np, wit_chunk(*out_f), 64, sc, &c, &app));
    int rc = finish_split(c, v, d_na, d_np, d_flag, S(stream));
    if (!rc) rc = launch_copy_u64(&c->meta->status, d_status, S(stream));
    if (rc) { ms_col_free(c); return rc; }
    *out = c;
    return MS_OK;
}

// ------------------------------------------------------------------ agg_sum
__global__ void k_agg_meta(ms_dev_meta *m, uint64_t compr)
{
    m->count = 1;
    m->size_used = 8;
    m->size_compr = compr;
}

extern "C" int ms_agg_sum(ms_col **out, const ms_col *in, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !in) return ms_fail(MS_ERR_INVALID, "bad argument");
    ms_col *c;
    MS_TRY(new_col(uncompr_desc(), 1, 64, S(stream), &c));
    ms_pred all = make_pred(MS_CMP_GE, 0);
    int rc = launch_select_sum(in, all, (uint64_t *)c->data, false, S(stream));
    if (!rc) {
        // agg_sum_compr.h:148-150 vs agg_sum_uncompr.h:126 (sizeCompr 8 vs 0)
        hipLaunchKernelGGL(k_agg_meta, dim3(1), dim3(1), 0, S(stream), c->meta, in->f.kind == MS_UNCOMPR ? 0ull : 8ull);
        if (hipGetLastError() != hipSuccess) rc = ms_fail(MS_ERR_HIP, "agg_sum meta");
    }
    if (rc) { ms_col_free(c); return rc; }
    *out = c;
    return MS_OK;
}

// ------------------------------------------------------------------ group, join
extern "C" int ms_group(ms_col **out_gr, ms_col **out_ext, const ms_format_desc *gr_f, const ms_format_desc *ext_f,
                        const ms_col *in_gr, const ms_col *in_data, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out_gr || !out_ext || !in_data) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(gr_f) || !wit_ok(gr_f) || !valid_format(ext_f) || !wit_ok(ext_f))
        return ms_fail(MS_ERR_INVALID, "invalid output format");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "group synchronises: not in a plan");
    uint64_t n, ng;
    MS_TRY(host_count(in_data, stream, &n));
    if (in_gr) {
        MS_TRY(host_count(in_gr, stream, &ng));
        if (ng != n)
            return ms_fail(MS_ERR_SIZE, "binary group: inGrCol and inDataCol must contain the same number of data "
                                        "elements");
    }
    temps sc(S(stream));
    const uint64_t *x, *g = nullptr, *d_n, *d_ng;
    MS_TRY(values_of(in_data, sc, &x, &d_n));
    if (in_gr) MS_TRY(values_of(in_gr, sc, &g, &d_ng));
    uint64_t *og = (uint64_t *)sc.get(8 * n + 64), *oe = (uint64_t *)sc.get(8 * n + 64);
    uint64_t *d_cnt = (uint64_t *)sc.get(64);
    if (!og || !oe || !d_cnt) return ms_fail(MS_ERR_NOMEM, "group scratch");
    MS_TRY(launch_group(x, g, n, og, oe, d_cnt, sc));
    ms_col *cg, *ce;
    uint64_t *app;
    MS_TRY(emit_new(*gr_f, og, d_n, n, wit_chunk(*gr_f), 0, sc, &cg, &app));
    int rc = emit_new(*ext_f, oe, d_cnt, n, wit_chunk(*ext_f), 0, sc, &ce, &app);
    if (rc) { ms_col_free(cg); return rc; }
    *out_gr = cg;
    *out_ext = ce;
    return MS_OK;
}

extern "C" int ms_join(ms_col **out_l, ms_col **out_r, const ms_format_desc *l_f, const ms_format_desc *r_f,
                       const ms_col *l, const ms_col *r, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out_l || !out_r || !l || !r) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(l_f) || !wit_ok(l_f) || !valid_format(r_f) || !wit_ok(r_f))
        return ms_fail(MS_ERR_INVALID, "invalid output format");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "join synchronises: not in a plan");
    uint64_t na, nb;
    MS_TRY(host_count(l, stream, &na));
    MS_TRY(host_count(r, stream, &nb));
    temps sc(S(stream));
    const uint64_t *a, *b, *d_na, *d_nb;
    MS_TRY(values_of(l, sc, &a, &d_na));
    MS_TRY(values_of(r, sc, &b, &d_nb));
    // the outer loop runs over the longer input (scalar/join_uncompr.h:53-66)
    const bool swap = na < nb;
    uint64_t *d_total = (uint64_t *)sc.get(64);
    if (!d_total) return ms_fail(MS_ERR_NOMEM, "join scratch");
    uint64_t *po = nullptr, *pi = nullptr, total = 0;
    MS_TRY(launch_join(swap ? b : a, swap ? nb : na, swap ? a : b, swap ? na : nb, false, &po, &pi, d_total, &total,
                       sc));
    const uint64_t *pl = swap ? pi : po, *pr = swap ? po : pi;
    ms_col *cl, *cr;
    uint64_t *app;
    MS_TRY(emit_new(*l_f, pl, d_total, total, wit_chunk(*l_f), 0, sc, &cl, &app));
    int rc = emit_new(*r_f, pr, d_total, total, wit_chunk(*r_f), 0, sc, &cr, &app);
    if (rc) { ms_col_free(cl); return rc; }
    *out_l = cl;
    *out_r = cr;
    return MS_OK;
}

extern "C" int ms_semi_join(ms_col **out, const ms_format_desc *out_f, const ms_col *l, const ms_col *r,
                            void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !l || !r) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    if (plan_capturing()) return ms_fail(MS_ERR_INVALID, "join synchronises: not in a plan");
    uint64_t na, nb;
    MS_TRY(host_count(l, stream, &na));
    MS_TRY(host_count(r, stream, &nb));
    temps sc(S(stream));
    const uint64_t *a, *b, *d_na, *d_nb;
    MS_TRY(values_of(l, sc, &a, &d_na));
    MS_TRY(values_of(r, sc, &b, &d_nb));
    uint64_t *d_total = (uint64_t *)sc.get(64);
    if (!d_total) return ms_fail(MS_ERR_NOMEM, "semi join scratch");
    uint64_t *po = nullptr, *pi = nullptr, total = 0;
    MS_TRY(launch_join(a, na, b, nb, true, &po, &pi, d_total, &total, sc));
    if (!po) po = d_total;   // no left value: nothing is read
    uint64_t *app;
    return emit_new(*out_f, po, d_total, na, wit_chunk(*out_f), 0, sc, out, &app);
}

// ------------------------------------------------------------------ set ops
extern "C" int ms_intersect_sorted(ms_col **out, const ms_format_desc *out_f, const ms_col *l, const ms_col *r,
                                   void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !l || !r) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    temps sc(S(stream));
    const uint64_t *a, *b, *d_na, *d_nb;
    MS_TRY(values_of(l, sc, &a, &d_na));
    MS_TRY(values_of(r, sc, &b, &d_nb));
    const uint64_t bound = l->count_bound;
    uint64_t *o = (uint64_t *)sc.get(bound * 8 + 64);
    uint64_t *d_cnt = (uint64_t *)sc.get(64);
    if (!o || !d_cnt) return ms_fail(MS_ERR_NOMEM, "intersect scratch");
    MS_TRY(launch_intersect(a, d_na, bound, b, d_nb, r->count_bound, o, d_cnt, sc));
    uint64_t *app;
    return emit_new(*out_f, o, d_cnt, bound, wit_chunk(*out_f), 0, sc, out, &app);
}

extern "C" int ms_merge_sorted(ms_col **out, const ms_format_desc *out_f, const ms_col *l, const ms_col *r,
                               void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !l || !r) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    temps sc(S(stream));
    const uint64_t *a, *b, *d_na, *d_nb;
    MS_TRY(values_of(l, sc, &a, &d_na));
    MS_TRY(values_of(r, sc, &b, &d_nb));
    const uint64_t bound = l->count_bound + r->count_bound;
    uint64_t *o = (uint64_t *)sc.get(bound * 8 + 64);
    uint64_t *d_cnt = (uint64_t *)sc.get(64);
    if (!o || !d_cnt) return ms_fail(MS_ERR_NOMEM, "merge scratch");
    MS_TRY(launch_merge(a, d_na, l->count_bound, b, d_nb, r->count_bound, o, d_cnt, sc));
    uint64_t *app;
    return emit_new(*out_f, o, d_cnt, bound, wit_chunk(*out_f), 0, sc, out, &app);
}

// The count calc_binary works on: the operands' common count, or 0 and
// MS_ERR_SIZE when they differ (scalar/calc_uncompr.h:84-88).
__global__ void k_calc_count(const ms_dev_meta *l, const ms_dev_meta *r, uint64_t *d_n, uint64_t *d_status)
{
    const uint64_t a = l->count, b = r->count;
    *d_n = a == b ? a : 0;
    *d_status = a == b ? 0 : (uint64_t)MS_ERR_SIZE;
}

__global__ void k_keep_status(uint64_t *dst, const uint64_t *src)
{
    if (*src) *dst = *src;
}

extern "C" int ms_calc_binary(ms_col **out, uint32_t op, const ms_format_desc *out_f, const ms_col *l,
                              const ms_col *r, void *stream)
{
    MS_TRY(ms_need_init());
    if (!out || !l || !r || op > MS_OP_XOR) return ms_fail(MS_ERR_INVALID, "bad argument");
    if (!valid_format(out_f) || !wit_ok(out_f)) return ms_fail(MS_ERR_INVALID, "invalid output format");
    if (!plan_capturing()) {
        // eagerly the mismatch is an error of this call, as in the reference
        ms_col_meta ml, mr;
        MS_TRY(ms_col_get_meta(l, &ml, stream));
        MS_TRY(ms_col_get_meta(r, &mr, stream));
        if (ml.count != mr.count)
            return ms_fail(MS_ERR_SIZE, "calc: inDataLCol and inDataRCol must contain the same number of data elements");
    }
    // inside a plan the counts are compared on the device: a mismatch
    // leaves an empty output whose status reads MS_ERR_SIZE (ms_col_check)
    temps sc(S(stream));
    const uint64_t *a, *b, *d_na, *d_nb;
    MS_TRY(values_of(l, sc, &a, &d_na));
    MS_TRY(values_of(r, sc, &b, &d_nb));
    uint64_t *d = (uint64_t *)sc.get(64);
    if (!d) return ms_fail(MS_ERR_NOMEM, "calc scratch");
    hipLaunchKernelGGL(k_calc_count, dim3(1), dim3(1), 0, S(stream), (const ms_dev_meta *)l->meta,
                       (const ms_dev_meta *)r->meta, d, d + 1);
    MS_HIP(hipGetLastError());
    const uint64_t bound = std::min(l->count_bound, r->count_bound);
    uint64_t *o = (uint64_t *)sc.get(bound * 8 + 64);
    if (!o) return ms_fail(MS_ERR_NOMEM, "calc scratch");
    MS_TRY(launch_calc(op, a, b, d, bound, o, S(stream)));
    uint64_t *app;
    ms_col *c;
    MS_TRY(emit_new(*out_f, o, d, bound, wit_chunk(*out_f), 0, sc, &c, &app));
    hipLaunchKernelGGL(k_keep_status, dim3(1), dim3(1), 0, S(stream), &c->meta->status, (const uint64_t *)(d + 1));
    if (hipGetLastError() != hipSuccess) {
        ms_col_free(c);
        return ms_fail(MS_ERR_HIP, "calc status");
    }
    *out = c;
    return MS_OK;
}
