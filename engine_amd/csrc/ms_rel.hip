// ms_rel.hip -- group and join over columns of any format (their values
// are decompressed first when compressed), results through the
// write-iterators of the requested output formats:
//   * group   (interfaces/group.h:38-99, scalar/group_uncompr.h:44-113):
//     group ids numbered in order of first appearance, and the position
//     of each group's first appearance (its extent);
//   * join    (interfaces/join.h:38-75, scalar/join_uncompr.h:44-96):
//     every equal pair of positions, ordered as the reference's nested
//     loops are: the outer loop runs over the longer input (the left one
//     when both are equally long), both ascending;
//   * semi join (interfaces/join.h:77-110, scalar/join_uncompr.h:98-128):
//     ascending left positions whose value occurs on the right.
// Design: a stable LSD radix sort of (value, position) pairs (hipCUB over
// rocPRIM) puts equal keys next to each other with positions ascending;
// group ids then come from two scans (segment heads in sorted order, first
// appearances in position order), joins from a binary search of each outer
// value in the sorted inner side.  All three read the inputs' counts on the
// host (the output sizes depend on them) and so synchronise.
#include <hipcub/hipcub.hpp>

#include "ms_units.hpp"

namespace ms {

__global__ void k_iota(uint64_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = i;
}

__global__ void k_gather_u64(const uint64_t *src, const uint64_t *idx, uint64_t n, uint64_t *dst)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

// Segment heads of the sorted keys (k1, and k2 when not null): head[k] = 1
// where a new key starts; first[pos[k]] = head[k] marks the first
// appearance of each key in position order (pos is a permutation).
__global__ void k_group_heads(const uint64_t *k1, const uint64_t *k2, const uint64_t *pos, uint64_t n, uint64_t *head,
                              uint64_t *first)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const bool h = k == 0 || k1[k] != k1[k - 1] || (k2 && k2[k] != k2[k - 1]);
        head[k] = h;
        first[pos[k]] = h;
    }
}

// Group id of every segment: the number of first appearances before its
// first position.
__global__ void k_group_seg_ids(const uint64_t *head, const uint64_t *seg_ex, const uint64_t *pos,
                                const uint64_t *gid_at, uint64_t n, uint64_t *seg_gid)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        if (head[k]) seg_gid[seg_ex[k]] = gid_at[pos[k]];
}

__global__ void k_group_out(const uint64_t *head, const uint64_t *seg_ex, const uint64_t *pos,
                            const uint64_t *seg_gid, const uint64_t *first, const uint64_t *gid_at, uint64_t n,
                            uint64_t *out_gr, uint64_t *out_ext, uint64_t *d_ngroups)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        out_gr[pos[k]] = seg_gid[seg_ex[k] + head[k] - 1];
        if (first[k]) out_ext[gid_at[k]] = k;   // k as a position here
        if (k == n - 1) *d_ngroups = gid_at[k] + first[k];
    }
}

__device__ __forceinline__ uint64_t lower_bound_u64(const uint64_t *a, uint64_t n, uint64_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// For every outer value: where its equal keys start in the sorted inner
// keys, and how many there are.
__global__ void k_join_count(const uint64_t *outer, uint64_t na, const uint64_t *keys, uint64_t nb, uint64_t *lo,
                             uint64_t *cnt)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = outer[i];
        const uint64_t l = lower_bound_u64(keys, nb, v);
        uint64_t h = l;
        if (l < nb && keys[l] == v) h = v == ~0ull ? nb : lower_bound_u64(keys, nb, v + 1);
        lo[i] = l;
        cnt[i] = h - l;
    }
}

__global__ void k_join_write(const uint64_t *lo, const uint64_t *cnt, const uint64_t *off, uint64_t na,
                             const uint64_t *inner_pos, uint64_t *out_outer, uint64_t *out_inner)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = off[i], l = lo[i];
        for (uint64_t k = 0; k < cnt[i]; k++) {
            out_outer[o + k] = i;
            out_inner[o + k] = inner_pos[l + k];
        }
    }
}

__global__ void k_semi_write(const uint64_t *cnt, const uint64_t *off, uint64_t na, uint64_t *out, uint64_t *d_n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na; i += (uint64_t)gridDim.x * blockDim.x) {
        if (cnt[i]) out[off[i]] = i;
        if (i == na - 1) *d_n = off[i] + (cnt[i] != 0);
    }
}

__global__ void k_flags_of(const uint64_t *cnt, uint64_t n, uint64_t *flag)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = cnt[i] != 0;
}

static uint32_t grid_n(uint64_t n) { return grid_for(n, WG * 4ull, 4096); }

// Stable ascending sort of n (key, value) pairs; vals_in may be null (keys
// only).
static int sort_pairs(const uint64_t *keys_in, uint64_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out,
                      uint64_t n, temps &sc)
{
    if (n == 0) return MS_OK;
    if (n >= (1ull << 31)) return ms_fail(MS_ERR_CAPACITY, "radix sort of %llu items", (unsigned long long)n);
    size_t bytes = 0;
    hipError_t e = vals_in ? hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                                                (int)n, 0, 64, sc.s)
                           : hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys_in, keys_out, (int)n, 0, 64, sc.s);
    if (e != hipSuccess) return ms_fail(MS_ERR_HIP, "radix sort size: %s", hipGetErrorString(e));
    void *tmp = sc.get(bytes + 64);
    if (!tmp) return ms_fail(MS_ERR_NOMEM, "radix sort scratch");
    e = vals_in ? hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0, 64,
                                                     sc.s)
                : hipcub::DeviceRadixSort::SortKeys(tmp, bytes, keys_in, keys_out, (int)n, 0, 64, sc.s);
    if (e != hipSuccess) return ms_fail(MS_ERR_HIP, "radix sort: %s", hipGetErrorString(e));
    return MS_OK;
}

int launch_group(const uint64_t *data, const uint64_t *gr, uint64_t n, uint64_t *out_gr, uint64_t *out_ext,
                 uint64_t *d_ngroups, temps &sc)
{
    if (n == 0) return launch_fill_u64(d_ngroups, 1, 0, sc.s);
    uint64_t *iota = (uint64_t *)sc.get(8 * n), *k1 = (uint64_t *)sc.get(8 * n), *p1 = (uint64_t *)sc.get(8 * n);
    uint64_t *head = (uint64_t *)sc.get(8 * n), *first = (uint64_t *)sc.get(8 * n);
    uint64_t *seg_ex = (uint64_t *)sc.get(8 * n), *gid_at = (uint64_t *)sc.get(8 * n);
    uint64_t *seg_gid = (uint64_t *)sc.get(8 * n);
    if (!iota || !k1 || !p1 || !head || !first || !seg_ex || !gid_at || !seg_gid)
        return ms_fail(MS_ERR_NOMEM, "group scratch");
    hipLaunchKernelGGL(k_iota, dim3(grid_n(n)), dim3(WG), 0, sc.s, iota, n);
    MS_TRY(sort_pairs(data, k1, iota, p1, n, sc));
    const uint64_t *ks = k1, *ks2 = nullptr, *ps = p1;
    if (gr) {
        // stable by the data value, then stable by the group id: sorted by
        // (group id, value) with positions ascending within equal pairs
        uint64_t *g1 = (uint64_t *)sc.get(8 * n), *k2 = (uint64_t *)sc.get(8 * n), *p2 = (uint64_t *)sc.get(8 * n);
        uint64_t *v2 = (uint64_t *)sc.get(8 * n);
        if (!g1 || !k2 || !p2 || !v2) return ms_fail(MS_ERR_NOMEM, "group scratch");
        hipLaunchKernelGGL(k_gather_u64, dim3(grid_n(n)), dim3(WG), 0, sc.s, gr, p1, n, g1);
        MS_TRY(sort_pairs(g1, k2, p1, p2, n, sc));
        hipLaunchKernelGGL(k_gather_u64, dim3(grid_n(n)), dim3(WG), 0, sc.s, data, p2, n, v2);
        ks = v2;
        ks2 = k2;
        ps = p2;
    }
    hipLaunchKernelGGL(k_group_heads, dim3(grid_n(n)), dim3(WG), 0, sc.s, ks, ks2, ps, n, head, first);
    MS_TRY(launch_exclusive_scan(head, seg_ex, n, sc));
    MS_TRY(launch_exclusive_scan(first, gid_at, n, sc));
    hipLaunchKernelGGL(k_group_seg_ids, dim3(grid_n(n)), dim3(WG), 0, sc.s, head, seg_ex, ps, gid_at, n, seg_gid);
    hipLaunchKernelGGL(k_group_out, dim3(grid_n(n)), dim3(WG), 0, sc.s, head, seg_ex, ps, seg_gid, first, gid_at, n,
                       out_gr, out_ext, d_ngroups);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

// Pairs of the nested loop join of outer (na values) with inner (nb):
// *total receives their number (synchronising the stream); the caller's
// outputs are allocated by `alloc` once it is known.
int launch_join(const uint64_t *outer, uint64_t na, const uint64_t *inner, uint64_t nb, bool semi,
                uint64_t **out_outer, uint64_t **out_inner, uint64_t *d_total, uint64_t *total, temps &sc)
{
    *total = 0;
    if (na == 0 || nb == 0) return launch_fill_u64(d_total, 1, 0, sc.s);
    uint64_t *iota = (uint64_t *)sc.get(8 * nb), *bk = (uint64_t *)sc.get(8 * nb), *bp = (uint64_t *)sc.get(8 * nb);
    uint64_t *lo = (uint64_t *)sc.get(8 * na), *cnt = (uint64_t *)sc.get(8 * na), *off = (uint64_t *)sc.get(8 * na);
    if (!iota || !bk || !bp || !lo || !cnt || !off) return ms_fail(MS_ERR_NOMEM, "join scratch");
    if (semi) {
        MS_TRY(sort_pairs(inner, bk, nullptr, nullptr, nb, sc));
    } else {
        hipLaunchKernelGGL(k_iota, dim3(grid_n(nb)), dim3(WG), 0, sc.s, iota, nb);
        MS_TRY(sort_pairs(inner, bk, iota, bp, nb, sc));
    }
    hipLaunchKernelGGL(k_join_count, dim3(grid_n(na)), dim3(WG), 0, sc.s, outer, na, bk, nb, lo, cnt);
    MS_HIP(hipGetLastError());
    if (semi) {
        uint64_t *flag = lo;   // the starts are not needed any more
        hipLaunchKernelGGL(k_flags_of, dim3(grid_n(na)), dim3(WG), 0, sc.s, cnt, na, flag);
        MS_TRY(launch_exclusive_scan(flag, off, na, sc));
        uint64_t *o = (uint64_t *)sc.get(8 * na + 64);
        if (!o) return ms_fail(MS_ERR_NOMEM, "semi join scratch");
        hipLaunchKernelGGL(k_semi_write, dim3(grid_n(na)), dim3(WG), 0, sc.s, cnt, off, na, o, d_total);
        MS_HIP(hipGetLastError());
        MS_HIP(hipMemcpyAsync(total, d_total, 8, hipMemcpyDeviceToHost, sc.s));
        MS_HIP(hipStreamSynchronize(sc.s));
        *out_outer = o;
        return MS_OK;
    }
    MS_TRY(launch_exclusive_scan(cnt, off, na, sc));
    uint64_t t[2];
    MS_HIP(hipMemcpyAsync(&t[0], off + na - 1, 8, hipMemcpyDeviceToHost, sc.s));
    MS_HIP(hipMemcpyAsync(&t[1], cnt + na - 1, 8, hipMemcpyDeviceToHost, sc.s));
    MS_HIP(hipStreamSynchronize(sc.s));
    *total = t[0] + t[1];
    MS_TRY(launch_fill_u64(d_total, 1, *total, sc.s));
    uint64_t *oa = (uint64_t *)sc.get(8 * *total + 64), *ob = (uint64_t *)sc.get(8 * *total + 64);
    if (!oa || !ob) return ms_fail(MS_ERR_NOMEM, "join output of %llu pairs", (unsigned long long)*total);
    hipLaunchKernelGGL(k_join_write, dim3(grid_n(na)), dim3(WG), 0, sc.s, lo, cnt, off, na, bp, oa, ob);
    MS_HIP(hipGetLastError());
    *out_outer = oa;
    *out_inner = ob;
    return MS_OK;
}

}  // namespace ms
