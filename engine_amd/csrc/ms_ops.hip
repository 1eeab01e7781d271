// ms_ops.hip -- project's gather, intersect_sorted, merge_sorted and
// calc_binary on gfx950.
//
//   gather     : out_i = data[pos_i] with random access into the compressed
//                part (static_vbp bit extraction, indexed dynamic_vbp /
//                for_f blocks) or the uncompressed rest
//                (project_compr.h:44-67, vbp_format_rndaccess.h:94-638).
//   intersect  : multiset intersection of ascending inputs, an element of
//                the left input survives if its rank among equal left
//                values is below the right input's count of that value
//                (scalar/intersect_uncompr.h:69-80 emits min multiplicity).
//   merge      : multiset union, equal pairs once (max multiplicity,
//                scalar/merge_uncompr.h:69-96), placed by merge-path ranks.
//   calc       : element-wise op mod 2^64 (scalar/calc_uncompr.h:98-100).
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

__device__ __forceinline__ uint64_t lower_bound(const uint64_t *a, uint64_t n, uint64_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t upper_bound(const uint64_t *a, uint64_t n, uint64_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t gather_one(const ms_colv &c, const ms_parts &pt, uint64_t p)
{
    if (p >= pt.n_compr) return ((const uint64_t *)(c.data + pt.rest_off))[p - pt.n_compr];
    if (c.g.kind == MS_STATIC_VBP) {
        const uint32_t S = c.g.lanes;
        return vbp_read((const uint64_t *)c.data, S, (uint32_t)(p % S), p / S, c.g.bw);
    }
    const uint64_t b = p / c.g.inner_bs;
    const uint64_t r = p - b * c.g.inner_bs;
    const uint64_t e = c.index[b];
    const uint32_t S = c.g.lanes;
    uint64_t v = vbp_read((const uint64_t *)(c.data + (e >> 8)), S, (uint32_t)(r % S), r / S, (uint32_t)(e & 0xff));
    if (c.g.kind == MS_FOR) v += c.refs[b / (c.g.outer_block / c.g.inner_bs)];
    return v;
}

__global__ __launch_bounds__(WG) void k_gather(ms_colv c, const uint64_t *pos, const uint64_t *d_np, uint64_t *out,
                                               uint64_t *status)
{
    const ms_parts pt = ms_parts_of(c.g, c.meta->count, c.meta->size_used, c.meta->size_compr);
    const uint64_t np = *d_np;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < np; i += (uint64_t)gridDim.x * WG) {
        const uint64_t p = pos[i];
        if (p >= pt.n) {
            atomicCAS((unsigned long long *)status, 0ull, (unsigned long long)MS_ERR_RANGE);
            out[i] = 0;
            continue;
        }
        out[i] = gather_one(c, pt, p);
    }
}

int launch_gather(const ms_col *data, const uint64_t *d_pos, const uint64_t *d_np, uint64_t np_bound,
                  uint64_t *d_out, uint64_t *d_status, temps &sc)
{
    ms_colv v;
    MS_TRY(colv_of(data, v));
    // gather_one reads uncompr_f, static_vbp_f and the block index of
    // dynamic_vbp_f / for_f; every other layout has no random access
    const uint32_t k = v.g.kind;
    if (k != MS_UNCOMPR && k != MS_STATIC_VBP && k != MS_DYNAMIC_VBP && k != MS_FOR)
        return ms_fail(MS_ERR_INVALID, "gather: no random access into format kind %u", k);
    if ((k == MS_DYNAMIC_VBP || k == MS_FOR) && !v.index)
        return ms_fail(MS_ERR_INVALID, "gather: column without a block index");
    hipLaunchKernelGGL(k_gather, dim3(grid_for(np_bound, WG * 4ull)), dim3(WG), 0, sc.s, v, d_pos, d_np, d_out,
                       d_status);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------ intersect
// Intersection flags (multiset minimum, scalar/intersect_uncompr.h:69-80):
// a[i] is kept when its rank among equal values of a is below the number
// of equal values in b.
// The same flags, a workgroup per chunk of ISECT_CHUNK values of a: the
// part of b that the chunk's value range can meet is staged in LDS when
// it fits, so the per-value searches in b run in LDS.
constexpr uint32_t ISECT_CHUNK = 512;
constexpr uint32_t ISECT_LDS = 2048;   // b values staged per workgroup (16 KiB: 9 workgroups per CU)

template <class Arr>
__device__ __forceinline__ uint64_t ub_in(const Arr &a, uint64_t n, uint64_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// b's range [lower_bound(first), upper_bound(last)) for every chunk of a,
// one thread per chunk (all searches in flight at once); d_nch = chunks.
__global__ __launch_bounds__(WG) void k_isect_ranges(const uint64_t *a, const uint64_t *d_na, const uint64_t *b,
                                                     const uint64_t *d_nb, uint64_t *rng, uint64_t *d_nch)
{
    const uint64_t na = *d_na, nb = *d_nb;
    const uint64_t nchunks = (na + ISECT_CHUNK - 1) / ISECT_CHUNK;
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_nch = nchunks;
    for (uint64_t ck = (uint64_t)blockIdx.x * WG + threadIdx.x; ck < nchunks; ck += (uint64_t)gridDim.x * WG) {
        const uint64_t i0 = ck * ISECT_CHUNK, i1 = i0 + ISECT_CHUNK < na ? i0 + ISECT_CHUNK : na;
        rng[2 * ck] = lower_bound(b, nb, a[i0]);
        rng[2 * ck + 1] = upper_bound(b, nb, a[i1 - 1]);
    }
}

// The flags of chunk ck: cnt[ck] = the chunk's kept values, and the kept
// values themselves, in order, at kept[ck * ISECT_CHUNK ..] (compacted in
// the chunk from its flag bitmask: bit k of word w is a[i0 + 64 w + k]'s
// flag).
constexpr uint32_t ISECT_WORDS = ISECT_CHUNK / 64;

__global__ __launch_bounds__(WG) void k_isect_mark(const uint64_t *a, const uint64_t *d_na, const uint64_t *b,
                                                   const uint64_t *d_nb, const uint64_t *rng, uint64_t *kept,
                                                   uint64_t *cnt)
{
    __shared__ uint64_t sb[ISECT_LDS];
    __shared__ uint32_t s_cnt[ISECT_WORDS];
    __shared__ uint64_t s_mask[ISECT_WORDS];
    const uint64_t na = *d_na, nb = *d_nb;
    const uint64_t nchunks = (na + ISECT_CHUNK - 1) / ISECT_CHUNK;
    for (uint64_t ck = blockIdx.x; ck < nchunks; ck += gridDim.x) {
        const uint64_t i0 = ck * ISECT_CHUNK, i1 = i0 + ISECT_CHUNK < na ? i0 + ISECT_CHUNK : na;
        const uint64_t b0 = rng[2 * ck], len = rng[2 * ck + 1] - b0;
        const bool staged = len <= ISECT_LDS;
        if (staged) {
            // all of a thread's loads in flight at once, then the LDS stores
            constexpr uint32_t SQ = ISECT_LDS / WG;
            uint64_t tmp[SQ];
#pragma unroll
            for (uint32_t q = 0; q < SQ; q++) {
                const uint64_t k = q * WG + threadIdx.x;
                tmp[q] = k < len ? b[b0 + k] : 0ull;
            }
#pragma unroll
            for (uint32_t q = 0; q < SQ; q++) {
                const uint64_t k = q * WG + threadIdx.x;
                if (k < len) sb[k] = tmp[q];
            }
        }
        constexpr uint32_t Q = ISECT_CHUNK / WG;   // values per thread
        uint64_t v[Q], prv[Q], lb[Q];
        bool in[Q];
#pragma unroll
        for (uint32_t q = 0; q < Q; q++) {
            const uint64_t i = i0 + q * WG + threadIdx.x;
            in[q] = i < i1;
            v[q] = in[q] ? a[i] : 0ull;
            prv[q] = in[q] && i > 0 ? a[i - 1] : 0ull;
            lb[q] = 0;
        }
        __syncthreads();
        if (staged && len) {
            // Q branch-free lower bounds in LDS side by side (same trip count)
            for (uint64_t step = 1ull << (63 - __builtin_clzll(len)); step; step >>= 1) {
#pragma unroll
                for (uint32_t q = 0; q < Q; q++) {
                    const uint64_t x = lb[q] + step;
                    if (x <= len && sb[x - 1] < v[q]) lb[q] = x;
                }
            }
        }
        uint32_t kbits = 0;   // bit q: the thread's value of row q is kept
#pragma unroll 1
        for (uint32_t q = 0; q < Q; q++) {
            const uint32_t e = q * WG + threadIdx.x;   // index in the chunk
            const uint64_t i = i0 + e;
            bool keep = false;
            if (in[q]) {
                const uint64_t x = v[q];
                const uint64_t rank = (i == 0 || prv[q] != x) ? 0 : i - lower_bound(a, i, x);
                uint64_t c = 0;
                if (staged) {
                    if (lb[q] < len && sb[lb[q]] == x) c = rank == 0 ? 1 : ub_in(sb, len, x) - lb[q];
                } else {
                    const uint64_t l = lower_bound(b, nb, x);
                    if (l < nb && b[l] == x) c = rank == 0 ? 1 : upper_bound(b, nb, x) - l;
                }
                keep = rank < c;
            }
            const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
            kbits |= (uint32_t)keep << q;
            if ((threadIdx.x & 63) == 0) {
                s_mask[e >> 6] = m;
                s_cnt[e >> 6] = (uint32_t)__builtin_popcountll(m);
            }
        }
        __syncthreads();
        // each kept value to its rank in the chunk: the kept values of the
        // rows before its row, then those before it in the row
        const uint32_t lane = threadIdx.x & 63;
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
#pragma unroll
        for (uint32_t q = 0; q < Q; q++) {
            const uint32_t row = (q * WG + threadIdx.x) >> 6;
            uint32_t pre = 0;
            for (uint32_t w = 0; w < row; w++) pre += s_cnt[w];
            if ((kbits >> q) & 1)
                kept[ck * ISECT_CHUNK + pre + (uint32_t)__builtin_popcountll(s_mask[row] & below)] = v[q];
        }
        if (threadIdx.x == 0) {
            uint64_t c = 0;
#pragma unroll
            for (uint32_t w = 0; w < ISECT_WORDS; w++) c += s_cnt[w];
            cnt[ck] = c;
        }
        __syncthreads();
    }
}

// Copies chunk ck's kept values to off[ck] on, a wave per chunk;
// *d_cnt = all kept values.
__global__ __launch_bounds__(WG) void k_isect_emit(const uint64_t *d_na, const uint64_t *kept, const uint64_t *cnt,
                                                   const uint64_t *off, uint64_t *out, uint64_t *d_cnt)
{
    const uint64_t na = *d_na;
    const uint64_t nchunks = (na + ISECT_CHUNK - 1) / ISECT_CHUNK;
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_cnt = nchunks ? off[nchunks - 1] + cnt[nchunks - 1] : 0;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t step = (uint64_t)gridDim.x * (WG / 64);
    for (uint64_t ck = (uint64_t)blockIdx.x * (WG / 64) + (threadIdx.x >> 6); ck < nchunks; ck += step) {
        const uint64_t base = off[ck], c = cnt[ck];
        for (uint64_t k = lane; k < c; k += 64) out[base + k] = kept[ck * ISECT_CHUNK + k];
    }
}

int launch_intersect(const uint64_t *a, const uint64_t *d_na, uint64_t na_bound, const uint64_t *b,
                     const uint64_t *d_nb, uint64_t nb_bound, uint64_t *d_out, uint64_t *d_cnt, temps &sc)
{
    (void)nb_bound;
    const uint64_t chunks = na_bound / ISECT_CHUNK + 1;
    uint64_t *rng = (uint64_t *)sc.get(chunks * 16 + 64);
    uint64_t *kept = (uint64_t *)sc.get(chunks * ISECT_CHUNK * 8 + 64);
    uint64_t *cnt = (uint64_t *)sc.get(chunks * 8 + 64);
    uint64_t *off = (uint64_t *)sc.get(chunks * 8 + 64);
    uint64_t *d_nch = (uint64_t *)sc.get(64);
    if (!rng || !kept || !cnt || !off || !d_nch) return ms_fail(MS_ERR_NOMEM, "intersect scratch");
    hipLaunchKernelGGL(k_isect_ranges, dim3(grid_for(chunks, WG, 2048)), dim3(WG), 0, sc.s, a, d_na, b, d_nb, rng,
                       d_nch);
    hipLaunchKernelGGL(k_isect_mark, dim3(grid_for(chunks, 1, 4096)), dim3(WG), 0, sc.s, a, d_na, b, d_nb,
                       (const uint64_t *)rng, kept, cnt);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan_n(cnt, off, d_nch, 0, chunks, sc));
    hipLaunchKernelGGL(k_isect_emit, dim3(grid_for(chunks, WG / 64, 8192)), dim3(WG), 0, sc.s, d_na,
                       (const uint64_t *)kept, (const uint64_t *)cnt, (const uint64_t *)off, d_out, d_cnt);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------ merge
__global__ __launch_bounds__(WG) void k_merge_rflags(const uint64_t *a, const uint64_t *d_na, const uint64_t *b,
                                                     const uint64_t *d_nb, uint64_t bound, uint64_t *flag)
{
    const uint64_t na = *d_na, nb = *d_nb;
    const uint64_t lim = nb + 1 < bound ? nb + 1 : bound;   // flag[nb] = 0 closes the scan
    for (uint64_t j = (uint64_t)blockIdx.x * WG + threadIdx.x; j < lim; j += (uint64_t)gridDim.x * WG) {
        uint64_t f = 0;
        if (j < nb) {
            const uint64_t v = b[j];
            const uint64_t rank = (j == 0 || b[j - 1] != v) ? 0 : j - lower_bound(b, j, v);
            const uint64_t lb = lower_bound(a, na, v);
            uint64_t cnt = 0;
            if (lb < na && a[lb] == v) cnt = upper_bound(a, na, v) - lb;
            f = rank >= cnt;
        }
        flag[j] = f;
    }
}

__global__ __launch_bounds__(WG) void k_merge_place(const uint64_t *a, const uint64_t *d_na, const uint64_t *b,
                                                    const uint64_t *d_nb, const uint64_t *flag, const uint64_t *er,
                                                    uint64_t *out, uint64_t *d_cnt)
{
    const uint64_t na = *d_na, nb = *d_nb;
    const uint64_t stride = (uint64_t)gridDim.x * WG;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < na; i += stride)
        out[i + er[lower_bound(b, nb, a[i])]] = a[i];
    for (uint64_t j = (uint64_t)blockIdx.x * WG + threadIdx.x; j < nb; j += stride)
        if (flag[j]) out[er[j] + upper_bound(a, na, b[j])] = b[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_cnt = na + er[nb];
}

int launch_merge(const uint64_t *a, const uint64_t *d_na, uint64_t na_bound, const uint64_t *b,
                 const uint64_t *d_nb, uint64_t nb_bound, uint64_t *d_out, uint64_t *d_cnt, temps &sc)
{
    const uint64_t bound = nb_bound + 1;
    uint64_t *flag = (uint64_t *)sc.get(bound * 8 + 64);
    uint64_t *er = (uint64_t *)sc.get(bound * 8 + 64);
    if (!flag || !er) return ms_fail(MS_ERR_NOMEM, "merge scratch");
    hipLaunchKernelGGL(k_merge_rflags, dim3(grid_for(bound, WG * 4ull)), dim3(WG), 0, sc.s, a, d_na, b, d_nb, bound,
                       flag);
    MS_HIP(hipGetLastError());
    MS_TRY(launch_exclusive_scan_n(flag, er, d_nb, 1, bound, sc));
    hipLaunchKernelGGL(k_merge_place, dim3(grid_for(na_bound + nb_bound + 1, WG * 4ull)), dim3(WG), 0, sc.s, a, d_na,
                       b, d_nb, flag, er, d_out, d_cnt);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// ------------------------------------------------------------------ calc
__device__ __forceinline__ uint64_t calc_op(uint32_t op, uint64_t x, uint64_t y)
{
    switch (op) {
    case MS_OP_ADD: return x + y;
    case MS_OP_SUB: return x - y;
    case MS_OP_MUL: return x * y;
    case MS_OP_DIV: return y ? x / y : 0;
    case MS_OP_MOD: return y ? x % y : 0;
    case MS_OP_MIN: return x < y ? x : y;
    case MS_OP_MAX: return x > y ? x : y;
    case MS_OP_AND: return x & y;
    case MS_OP_OR: return x | y;
    default: return x ^ y;
    }
}

__global__ __launch_bounds__(WG) void k_calc(uint32_t op, const uint64_t *a, const uint64_t *b, const uint64_t *d_n,
                                             uint64_t *out)
{
    const uint64_t n = *d_n;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        out[i] = calc_op(op, a[i], b[i]);
}

int launch_calc(uint32_t op, const uint64_t *a, const uint64_t *b, const uint64_t *d_n, uint64_t n_bound,
                uint64_t *d_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_calc, dim3(grid_for(n_bound, WG * 4ull)), dim3(WG), 0, s, op, a, b, d_n, d_out);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

__global__ void k_copy_u64(uint64_t *dst, const uint64_t *src) { *dst = *src; }

int launch_copy_u64(uint64_t *dst, const uint64_t *src, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(1), 0, s, dst, src);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// meta = {count, count * size_per_value, size_compr_is_used ? size_used : 0}
__global__ void k_set_meta(ms_dev_meta *m, const uint64_t *d_count, uint64_t per, uint64_t compr_is_used)
{
    const uint64_t c = *d_count;
    m->count = c;
    m->size_used = c * per;
    m->size_compr = compr_is_used ? c * per : 0;
}

int launch_set_meta(ms_dev_meta *m, const uint64_t *d_count, uint64_t size_per_value, uint64_t size_compr_is_used,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_set_meta, dim3(1), dim3(1), 0, s, m, d_count, size_per_value, size_compr_is_used);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

}  // namespace ms
