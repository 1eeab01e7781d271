// ms_learn.hip -- learned per-chunk format selection on gfx950.
//
// The reference picks formats from calibration tables and an exhaustive
// size search (format_selector_commons.h:25-103,
// exhaustive_size_format_selector.h:35-141, calibration/bw_prof.cpp).  Here:
//   * a chunk generator writes seeded synthetic chunks of 2^16 values on the
//     device (training data for the selector);
//   * a statistics kernel (one wave per chunk) computes per chunk the exact
//     integer statistics every candidate's size follows from: per-512-value
//     block bit widths of x, of the step-8 deltas and of the for_f offsets,
//     run count, order and range; from them the exact compressed size of
//     each candidate format, the 24 features of the network and, given a
//     per-(format, bit width) decode-cost table, a decode-time label;
//   * an FP32 MLP 24 -> 64 -> 64 -> 10 (ReLU) with forward, backward and
//     Adam kernels, trained on those labels, predicts log2 size and log2
//     decode time per candidate;
//   * a selection kernel picks per chunk the fastest predicted candidate
//     among those predicted within a tolerance of the smallest, and checks
//     the pick against the exact sizes (fallback: the exact smallest).
#include <math.h>

#include <vector>

#include "ms_units.hpp"

namespace ms {

constexpr uint32_t CHUNK = MS_CHUNK_VALUES;
constexpr uint32_t CBLOCKS = CHUNK / 512;

// ------------------------------------------------------------------ generator
struct gen_spec {
    uint32_t kind, a, b, table;   // table: index of the geometric table (sorted, rle)
    uint64_t seed, base;
};

__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Geometric variate on {1 .. MS_GEO_LEN}: 1 + #{k >= 1 : r < thr[k]}, thr
// decreasing (thr[0] = 2^64 - 1).
__device__ __forceinline__ uint64_t geo(const uint64_t *thr, uint64_t r)
{
    uint32_t lo = 0, hi = MS_GEO_LEN;   // invariant: thr[lo] > r (lo = 0 is the sentinel), thr[hi] <= r or hi = LEN
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (r < thr[mid]) lo = mid; else hi = mid;
    }
    return (uint64_t)lo + 1;
}

__device__ __forceinline__ uint64_t top_bits(uint64_t r, uint32_t bits) { return bits ? r >> (64 - bits) : 0ull; }

// One wave per chunk; value i = 512 b + 64 k + t.
__global__ __launch_bounds__(WG) void k_gen_chunks(const gen_spec *specs, uint64_t nchunks, const uint64_t *tables,
                                                   uint64_t *out)
{
    const uint64_t c = (uint64_t)blockIdx.x * WAVES + wave_index();
    if (c >= nchunks) return;
    const gen_spec sp = specs[c];
    const uint32_t t = threadIdx.x & 63;
    uint64_t *o = out + c * CHUNK;
    const uint64_t *thr = tables + (uint64_t)sp.table * MS_GEO_LEN;
    switch (sp.kind) {
    case MS_GEN_UNIFORM:
        for (uint32_t i = t; i < CHUNK; i += 64) o[i] = top_bits(splitmix(sp.seed, i), sp.a);
        break;
    case MS_GEN_DVBP:
        for (uint32_t b = 0; b < CBLOCKS; b++) {
            const uint32_t bw = sp.a + (uint32_t)(splitmix(sp.seed ^ 0xB1, b) % (sp.b - sp.a + 1));
            const uint32_t force = (uint32_t)(splitmix(sp.seed ^ 0xB2, b) % 512);
            for (uint32_t k = 0; k < 8; k++) {
                const uint32_t r = 64 * k + t, i = 512 * b + r;
                uint64_t x = top_bits(splitmix(sp.seed, i), bw);
                if (r == force) x |= 1ull << (bw - 1);
                o[i] = x;
            }
        }
        break;
    case MS_GEN_SORTED: {
        // x_0 = base, x_i = x_{i-1} + g_i with g_i geometric (i >= 1)
        uint64_t run = sp.base;
        for (uint32_t i0 = 0; i0 < CHUNK; i0 += 64) {
            const uint32_t i = i0 + t;
            uint64_t g = i == 0 ? 0ull : geo(thr, splitmix(sp.seed, i));
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint64_t v = __shfl_up(g, off, MS_WAVE);
                if (t >= off) g += v;
            }
            o[i] = run + g;
            run += __shfl(g, 63, MS_WAVE);
        }
        break;
    }
    case MS_GEN_FOR:
        for (uint32_t i = t; i < CHUNK; i += 64) o[i] = sp.base + top_bits(splitmix(sp.seed, i), sp.a);
        break;
    default: {   // MS_GEN_RLE: run k has length geo(seed^0x77, k) and value top b bits of (seed^0x78, k)
        uint64_t start = 0;
        for (uint64_t k0 = 0; start < CHUNK; k0 += 64) {
            const uint64_t k = k0 + t;
            const uint64_t len = geo(thr, splitmix(sp.seed ^ 0x77, k));
            const uint64_t val = top_bits(splitmix(sp.seed ^ 0x78, k), sp.b);
            uint64_t end = len;
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint64_t v = __shfl_up(end, off, MS_WAVE);
                if (t >= off) end += v;
            }
            const uint64_t s = start + end - len;
            for (uint64_t i = s; i < s + len && i < CHUNK; i++) o[i] = val;
            start += __shfl(end, 63, MS_WAVE);
        }
        break;
    }
    }
}

// ------------------------------------------------------------------ statistics
__device__ __forceinline__ uint32_t eff_bw(uint64_t x)   // math.h:108-113 (x > 0)
{
    return 64 - __clzll(x);
}

// One wave per chunk of 2^16 values; value i = 512 b + 64 k + t is in
// delta/for lane i % 8 = t % 8.  Integer statistics (MS_ST_*), exact.
__global__ __launch_bounds__(WG) void k_chunk_stats(const uint64_t *__restrict__ vals, uint64_t nchunks,
                                                    uint64_t *__restrict__ st, const float *__restrict__ cost,
                                                    float *__restrict__ tlab)
{
    const uint64_t c = (uint64_t)blockIdx.x * WAVES + wave_index();
    if (c >= nchunks) return;
    const uint32_t t = threadIdx.x & 63;
    const uint64_t *x = vals + c * CHUNK;
    uint64_t sbw = 0, sdbw = 0, sfbw = 0, sbw2 = 0, hist = 0;
    uint32_t mbw = 0, mdbw = 0, mfbw = 0, dless = 0, fless = 0;
    uint64_t runs = 0, nondec = 0, zeros = 0, mn = ~0ull, mx = 0;
    float tdv = 0.f, tde = 0.f, tfo = 0.f;
    uint64_t lastv7 = 0;   // x(t, 7) of the previous block
    for (uint32_t ob = 0; ob < CBLOCKS / 2; ob++) {
        uint64_t v[2][8];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int k = 0; k < 8; k++) v[h][k] = x[(2 * ob + h) * 512 + 64 * k + t];
        // for_f reference value of the 1024-value block (for.h:121-161):
        // signed minimum per lane (t % 8), unsigned minimum over the lanes
        int64_t sm = INT64_MAX;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int k = 0; k < 8; k++) sm = (int64_t)v[h][k] < sm ? (int64_t)v[h][k] : sm;
        for (uint32_t o = 32; o >= 8; o >>= 1) {
            const int64_t y = (int64_t)__shfl_xor((uint64_t)sm, o, MS_WAVE);
            sm = y < sm ? y : sm;
        }
        uint64_t ref = (uint64_t)sm;
        for (uint32_t o = 4; o >= 1; o >>= 1) {
            const uint64_t y = __shfl_xor(ref, o, MS_WAVE);
            ref = y < ref ? y : ref;
        }
        if (ref == ~0ull) ref -= 1;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint64_t orx = 0, ord = 0, orf = 0;
            const bool first = ob == 0 && h == 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t xv = v[h][k];
                // x_{i-8}: lane t-8 at the same k, else lane t+56 one k back
                const uint64_t up8 = __shfl(xv, (t - 8) & 63, MS_WAVE);
                const uint64_t pk = k ? v[h][k - 1] : (h ? v[0][7] : lastv7);
                const uint64_t back8 = __shfl(pk, (t + 56) & 63, MS_WAVE);
                const uint64_t p8 = t >= 8 ? up8 : ((first && k == 0) ? 0ull : back8);
                // x_{i-1}
                const uint64_t up1 = __shfl(xv, (t - 1) & 63, MS_WAVE);
                const uint64_t back1 = __shfl(pk, 63, MS_WAVE);
                const bool has_prev = !(first && k == 0 && t == 0);
                const uint64_t p1 = t >= 1 ? up1 : back1;
                orx |= xv;
                ord |= xv - p8;
                orf |= xv - ref;
                runs += !has_prev || xv != p1;
                nondec += has_prev && xv >= p1;
                zeros += xv == 0;
                mn = xv < mn ? xv : mn;
                mx = xv > mx ? xv : mx;
            }
            orx = wave_or_u64(orx);
            ord = wave_or_u64(ord);
            orf = wave_or_u64(orf);
            const uint32_t bw = eff_bw(orx | 1), dbw = eff_bw(ord | 1), fbw = eff_bw(orf | 1);
            sbw += bw; sdbw += dbw; sfbw += fbw; sbw2 += (uint64_t)bw * bw;
            mbw = bw > mbw ? bw : mbw;
            mdbw = dbw > mdbw ? dbw : mdbw;
            mfbw = fbw > mfbw ? fbw : mfbw;
            dless += dbw < bw;
            fless += fbw < bw;
            hist += 1ull << (8 * ((bw - 1) >> 3));
            if (cost) {
                tdv += cost[1 * 65 + bw];
                tde += cost[2 * 65 + dbw];
                tfo += cost[3 * 65 + fbw];
            }
        }
        lastv7 = v[1][7];
    }
    // per-lane counts and extremes over the wave
    runs = wave_sum_u64(runs);
    nondec = wave_sum_u64(nondec);
    zeros = wave_sum_u64(zeros);
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t a = __shfl_xor(mn, o, MS_WAVE), b = __shfl_xor(mx, o, MS_WAVE);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (t == 0) {
        uint64_t *s = st + c * MS_ST_COUNT;
        s[MS_ST_SUM_BW] = sbw;
        s[MS_ST_MAX_BW] = mbw;
        s[MS_ST_SUM_DBW] = sdbw;
        s[MS_ST_MAX_DBW] = mdbw;
        s[MS_ST_SUM_FBW] = sfbw;
        s[MS_ST_MAX_FBW] = mfbw;
        s[MS_ST_RUNS] = runs;
        s[MS_ST_NONDEC] = nondec;
        s[MS_ST_ZEROS] = zeros;
        s[MS_ST_MIN] = mn;
        s[MS_ST_MAX] = mx;
        s[MS_ST_SUM_BW2] = sbw2;
        s[MS_ST_DLESS] = dless;
        s[MS_ST_FLESS] = fless;
        s[MS_ST_HIST] = hist;
        s[MS_ST_RESERVED] = 0;
        if (tlab) {
            float *tl = tlab + c * MS_NCAND;
            tl[MS_CAND_STATIC] = CBLOCKS * cost[0 * 65 + mbw];
            tl[MS_CAND_DVBP] = tdv;
            tl[MS_CAND_DELTA] = tde;
            tl[MS_CAND_FOR] = tfo;
            tl[MS_CAND_RLE] = cost[4 * 65 + 0] + cost[4 * 65 + 1] * (float)runs;
        }
    }
}

}  // namespace ms

namespace ms {

// Features of a chunk from its statistics (MS_NFEAT floats) and the exact
// size in bytes of each candidate format, for 2^16-value chunks (no rest):
//   static_vbp<maxbw, 8>              : 2^16 * maxbw / 8
//   dynamic_vbp<512, 64, 8>           : two pages of 64 meta bytes + 64 bw per block
//   delta<1024, 8, dvbp<512, 64, 8>>  : per 1024 values one inner call (64 meta bytes) + 64 dbw per block
//   for<1024, 8, dvbp<512, 64, 8>>    : + one page of 8 references per 8 blocks of 1024
//   rle                               : 16 bytes per run
__global__ __launch_bounds__(WG) void k_chunk_features(const uint64_t *__restrict__ st, uint64_t nchunks,
                                                       float *__restrict__ feat, uint64_t *__restrict__ sizes)
{
    const uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (c >= nchunks) return;
    const uint64_t *s = st + c * MS_ST_COUNT;
    const float nb = (float)CBLOCKS;
    if (feat) {
        float *f = feat + c * MS_NFEAT;
        const uint64_t hist = s[MS_ST_HIST];
        for (int k = 0; k < 8; k++) f[k] = (float)((hist >> (8 * k)) & 0xff) / nb;
        f[8] = (float)s[MS_ST_SUM_BW] / nb / 64.f;
        f[9] = (float)s[MS_ST_MAX_BW] / 64.f;
        f[10] = (float)s[MS_ST_SUM_DBW] / nb / 64.f;
        f[11] = (float)s[MS_ST_MAX_DBW] / 64.f;
        f[12] = (float)s[MS_ST_SUM_FBW] / nb / 64.f;
        f[13] = (float)s[MS_ST_MAX_FBW] / 64.f;
        f[14] = log2f((float)s[MS_ST_RUNS]) / 16.f;
        f[15] = (float)s[MS_ST_RUNS] / (float)CHUNK;
        f[16] = (float)s[MS_ST_DLESS] / nb;
        f[17] = (float)s[MS_ST_FLESS] / nb;
        const uint64_t mn = s[MS_ST_MIN], mx = s[MS_ST_MAX];
        f[18] = (float)(mx == mn ? 0 : eff_bw(mx - mn)) / 64.f;
        f[19] = (float)(mn ? eff_bw(mn) : 0) / 64.f;
        f[20] = (float)(mx ? eff_bw(mx) : 0) / 64.f;
        f[21] = (float)s[MS_ST_NONDEC] / (float)(CHUNK - 1);
        const float mean = (float)s[MS_ST_SUM_BW] / nb;
        const float var = (float)s[MS_ST_SUM_BW2] / nb - mean * mean;
        f[22] = sqrtf(var > 0.f ? var : 0.f) / 32.f;
        f[23] = (float)s[MS_ST_ZEROS] / (float)CHUNK;
    }
    if (sizes) {
        uint64_t *z = sizes + c * MS_NCAND;
        z[MS_CAND_STATIC] = (uint64_t)CHUNK / 8 * s[MS_ST_MAX_BW];
        z[MS_CAND_DVBP] = (CBLOCKS / 64) * 64 + 64 * s[MS_ST_SUM_BW];
        z[MS_CAND_DELTA] = (CHUNK / 1024) * 64 + 64 * s[MS_ST_SUM_DBW];
        z[MS_CAND_FOR] = (CHUNK / 1024 / 8) * 64 + (CHUNK / 1024) * 64 + 64 * s[MS_ST_SUM_FBW];
        z[MS_CAND_RLE] = 16 * s[MS_ST_RUNS];
    }
}

__global__ void k_fill_f32(float *p, uint32_t n, float v)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ MLP
// Parameters, one float array: W1[IN][HID] b1[HID] W2[HID][HID] b2[HID]
// W3[HID][OUT] b3[OUT] (row = input unit).
constexpr int IN = MS_MLP_IN, HID = MS_MLP_HIDDEN, OUT = MS_MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + IN * HID, OW2 = OB1 + HID, OB2 = OW2 + HID * HID, OW3 = OB2 + HID,
              OB3 = OW3 + HID * OUT, NPARAM = OB3 + OUT;
static_assert(NPARAM == MS_MLP_NPARAM, "parameter count");
constexpr int SPW = 64;   // samples per workgroup

// Forward + backward over SPW samples per workgroup (4 threads per sample,
// 16 hidden units each).  Loss: sum over samples and outputs of
// w_o (y - t)^2 / norm.  Writes the workgroup's gradient and loss to
// grad_part[blockIdx.x] / loss_part[blockIdx.x].
__global__ __launch_bounds__(WG) void k_mlp_fwd_bwd(const float *__restrict__ prm, const float *__restrict__ X,
                                                    const float *__restrict__ Y, const uint32_t *__restrict__ order,
                                                    uint64_t first, uint64_t count, const float *__restrict__ wout,
                                                    float norm, float *__restrict__ grad_part,
                                                    float *__restrict__ loss_part)
{
    __shared__ float sP[NPARAM];
    __shared__ float sX[SPW][IN];
    __shared__ float sH1[SPW][HID];
    __shared__ float sH2[SPW][HID];
    __shared__ float sD[SPW][HID];     // dH2, then dH1
    __shared__ float sDY[SPW][OUT];
    __shared__ float sL[WG];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < NPARAM; i += WG) sP[i] = prm[i];
    const uint64_t s0 = first + (uint64_t)blockIdx.x * SPW;
    for (uint32_t i = tid; i < SPW * IN; i += WG) {
        const uint32_t s = i / IN, k = i % IN;
        const uint64_t g = s0 + s;
        sX[s][k] = g < first + count ? X[(uint64_t)(order ? order[g] : g) * IN + k] : 0.f;
    }
    __syncthreads();
    const uint32_t s = tid >> 2, q = tid & 3;
    const uint64_t gs = s0 + s;
    const bool live = gs < first + count;
    // layer 1
    for (int u = 0; u < 16; u++) {
        const int h = q * 16 + u;
        float a = sP[OB1 + h];
        for (int k = 0; k < IN; k++) a += sX[s][k] * sP[OW1 + k * HID + h];
        sH1[s][h] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int u = 0; u < 16; u++) {
        const int h = q * 16 + u;
        float a = sP[OB2 + h];
        for (int k = 0; k < HID; k++) a += sH1[s][k] * sP[OW2 + k * HID + h];
        sH2[s][h] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    float l = 0.f;
    for (int o = q; o < OUT; o += 4) {
        float a = sP[OB3 + o];
        for (int k = 0; k < HID; k++) a += sH2[s][k] * sP[OW3 + k * OUT + o];
        const float e = live ? a - Y[(uint64_t)(order ? order[gs] : gs) * OUT + o] : 0.f;
        l += wout[o] * e * e;
        sDY[s][o] = 2.f * wout[o] * e / norm;
    }
    sL[tid] = l;
    __syncthreads();
    // dH2 = dY W3^T masked by ReLU
    for (int u = 0; u < 16; u++) {
        const int h = q * 16 + u;
        float a = 0.f;
        for (int o = 0; o < OUT; o++) a += sDY[s][o] * sP[OW3 + h * OUT + o];
        sD[s][h] = sH2[s][h] > 0.f ? a : 0.f;
    }
    __syncthreads();
    float *gp = grad_part + (uint64_t)blockIdx.x * NPARAM;
    // dW3, db3
    for (uint32_t p = tid; p < HID * OUT + OUT; p += WG) {
        float a = 0.f;
        if (p < HID * OUT) {
            const uint32_t h = p / OUT, o = p % OUT;
            for (int r = 0; r < SPW; r++) a += sH2[r][h] * sDY[r][o];
        } else {
            const uint32_t o = p - HID * OUT;
            for (int r = 0; r < SPW; r++) a += sDY[r][o];
        }
        gp[OW3 + p] = a;
    }
    // dW2, db2
    for (uint32_t p = tid; p < HID * HID + HID; p += WG) {
        float a = 0.f;
        if (p < HID * HID) {
            const uint32_t i = p / HID, h = p % HID;
            for (int r = 0; r < SPW; r++) a += sH1[r][i] * sD[r][h];
        } else {
            const uint32_t h = p - HID * HID;
            for (int r = 0; r < SPW; r++) a += sD[r][h];
        }
        gp[OW2 + p] = a;
    }
    // dH1 = dH2 W2^T masked by ReLU, into sH2 once every thread is done
    // reading it (dW3 above)
    __syncthreads();
    for (int u = 0; u < 16; u++) {
        const int i = q * 16 + u;
        float a = 0.f;
        for (int h = 0; h < HID; h++) a += sD[s][h] * sP[OW2 + i * HID + h];
        sH2[s][i] = sH1[s][i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // dW1, db1
    for (uint32_t p = tid; p < IN * HID + HID; p += WG) {
        float a = 0.f;
        if (p < IN * HID) {
            const uint32_t k = p / HID, h = p % HID;
            for (int r = 0; r < SPW; r++) a += sX[r][k] * sH2[r][h];
        } else {
            const uint32_t h = p - IN * HID;
            for (int r = 0; r < SPW; r++) a += sH2[r][h];
        }
        gp[OW1 + p] = a;
    }
    if (tid == 0) {
        float a = 0.f;
        for (int i = 0; i < WG; i++) a += sL[i];
        loss_part[blockIdx.x] = a;
    }
}

// Sums the workgroup gradients; then either writes them (adam == 0) or
// applies one Adam step to the parameters.
__global__ __launch_bounds__(WG) void k_mlp_reduce(float *__restrict__ prm, float *__restrict__ m1,
                                                   float *__restrict__ m2, const float *__restrict__ grad_part,
                                                   uint32_t nparts, const float *__restrict__ loss_part,
                                                   float loss_scale, float *__restrict__ loss_acc, int adam,
                                                   float lr, float b1, float b2, float eps, float c1, float c2,
                                                   float *__restrict__ grad_out)
{
    const uint32_t p = blockIdx.x * WG + threadIdx.x;
    if (p == NPARAM && loss_acc) {
        float a = 0.f;
        for (uint32_t w = 0; w < nparts; w++) a += loss_part[w];
        *loss_acc += a * loss_scale;
    }
    if (p >= NPARAM) return;
    float g = 0.f;
    for (uint32_t w = 0; w < nparts; w++) g += grad_part[(uint64_t)w * NPARAM + p];
    if (!adam) {
        grad_out[p] = g;
        return;
    }
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    prm[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// Inference, one thread per sample (parameters read through the cache).
__global__ __launch_bounds__(WG) void k_mlp_predict(const float *__restrict__ prm, const float *__restrict__ X,
                                                    uint64_t n, float *__restrict__ Yp)
{
    __shared__ float sP[NPARAM];
    for (uint32_t i = threadIdx.x; i < NPARAM; i += WG) sP[i] = prm[i];
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (s >= n) return;
    float x[IN], h1[HID];
    for (int k = 0; k < IN; k++) x[k] = X[s * IN + k];
    for (int h = 0; h < HID; h++) {
        float a = sP[OB1 + h];
        for (int k = 0; k < IN; k++) a += x[k] * sP[OW1 + k * HID + h];
        h1[h] = a > 0.f ? a : 0.f;
    }
    float y[OUT];
    for (int o = 0; o < OUT; o++) y[o] = sP[OB3 + o];
    for (int h = 0; h < HID; h++) {
        float a = sP[OB2 + h];
        for (int k = 0; k < HID; k++) a += h1[k] * sP[OW2 + k * HID + h];
        a = a > 0.f ? a : 0.f;
        for (int o = 0; o < OUT; o++) y[o] += a * sP[OW3 + h * OUT + o];
    }
    for (int o = 0; o < OUT; o++) Yp[s * OUT + o] = y[o];
}

// Per chunk: the candidate with the smallest predicted size, unless another
// candidate predicted within size_tol of that size is predicted faster by
// more than time_margin (then the fastest such one); if the pick is more
// than exact_tol larger than the exact smallest size, the exact smallest
// instead (bit 7 of the choice marks such a fallback).  Predictions are
// (log2 bytes - size_off) / size_scale and (log2 ns - time_off) / time_scale.
__global__ __launch_bounds__(WG) void k_select_formats(const float *__restrict__ pred,
                                                       const uint64_t *__restrict__ sizes, uint64_t n,
                                                       float size_tol, float exact_tol, float time_margin,
                                                       float4 norm, uint32_t rule, uint8_t *__restrict__ choice)
{
    const uint64_t c = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (c >= n) return;
    const float *p = pred + c * OUT;
    float ps[MS_NCAND], pt[MS_NCAND];
    int small = 0;
    for (int f = 0; f < MS_NCAND; f++) {
        ps[f] = exp2f(p[f] * norm.y + norm.x);
        pt[f] = exp2f(p[MS_NCAND + f] * norm.w + norm.z);
        small = ps[f] < ps[small] ? f : small;
    }
    int pick = small;
    for (int f = 0; f < MS_NCAND; f++) {
        if (ps[f] > (1.f + size_tol) * ps[small]) continue;
        if (pt[f] * (1.f + time_margin) < pt[small] && pt[f] < pt[pick]) pick = f;
    }
    const uint64_t *z = sizes + c * MS_NCAND;
    int ex = 0;
    for (int f = 1; f < MS_NCAND; f++) ex = z[f] < z[ex] ? f : ex;
    if (rule == 1) {
        // exact sizes admit, predicted times choose
        int fast = ex;
        for (int f = 0; f < MS_NCAND; f++) {
            if ((double)z[f] > (1.0 + exact_tol) * (double)z[ex]) continue;
            if (pt[f] * (1.f + time_margin) < pt[ex] && pt[f] < pt[fast]) fast = f;
        }
        choice[c] = (uint8_t)(fast == ex ? (ex | 0x80) : fast);
        return;
    }
    uint8_t out = (uint8_t)pick;
    if ((double)z[pick] > (1.0 + exact_tol) * (double)z[ex]) out = (uint8_t)(ex | 0x80);
    choice[c] = out;
}

}  // namespace ms

using namespace ms;

// ------------------------------------------------------------------ C-ABI
extern "C" int ms_gen_chunks(uint64_t *d_out, const ms_chunk_spec *specs, uint64_t nchunks,
                             const uint64_t *geo_tables, uint32_t ntables, void *stream)
{
    MS_TRY(ms_need_init());
    if (!d_out || (!specs && nchunks) || (!geo_tables && ntables)) return ms_fail(MS_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    if (nchunks == 0) return MS_OK;
    for (uint64_t c = 0; c < nchunks; c++) {
        const ms_chunk_spec &p = specs[c];
        const bool bad = p.kind > MS_GEN_RLE || (p.kind == MS_GEN_UNIFORM && p.a > 64) ||
                         (p.kind == MS_GEN_DVBP && (p.a < 1 || p.b > 64 || p.a > p.b)) ||
                         (p.kind == MS_GEN_FOR && p.a > 64) || (p.kind == MS_GEN_RLE && p.b > 64) ||
                         ((p.kind == MS_GEN_SORTED || p.kind == MS_GEN_RLE) && p.table >= ntables);
        if (bad) return ms_fail(MS_ERR_INVALID, "chunk spec %llu out of range", (unsigned long long)c);
    }
    temps sc(s);
    gen_spec *d_specs = (gen_spec *)sc.get(nchunks * sizeof(gen_spec));
    uint64_t *d_tab = (uint64_t *)sc.get((uint64_t)(ntables ? ntables : 1) * MS_GEO_LEN * 8);
    if (!d_specs || !d_tab) return ms_fail(MS_ERR_NOMEM, "generator scratch");
    static_assert(sizeof(gen_spec) == sizeof(ms_chunk_spec), "spec layout");
    MS_HIP(hipMemcpyAsync(d_specs, specs, nchunks * sizeof(gen_spec), hipMemcpyHostToDevice, s));
    if (ntables)
        MS_HIP(hipMemcpyAsync(d_tab, geo_tables, (uint64_t)ntables * MS_GEO_LEN * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gen_chunks, dim3((nchunks + WAVES - 1) / WAVES), dim3(WG), 0, s, (const gen_spec *)d_specs,
                       nchunks, (const uint64_t *)d_tab, d_out);
    MS_HIP(hipGetLastError());
    // the host copies above read pageable memory that must stay valid
    MS_HIP(hipStreamSynchronize(s));
    return MS_OK;
}

extern "C" int ms_chunk_stats(const uint64_t *d_values, uint64_t nchunks, uint64_t *d_stats,
                              const float *cost_table, float *d_time_labels, void *stream)
{
    MS_TRY(ms_need_init());
    if (!d_values || !d_stats || (d_time_labels && !cost_table)) return ms_fail(MS_ERR_INVALID, "null argument");
    if (nchunks == 0) return MS_OK;
    hipStream_t s = (hipStream_t)stream;
    temps sc(s);
    float *d_cost = nullptr;
    if (cost_table) {
        d_cost = (float *)sc.get(MS_NCAND * 65 * sizeof(float));
        if (!d_cost) return ms_fail(MS_ERR_NOMEM, "cost table");
        MS_HIP(hipMemcpyAsync(d_cost, cost_table, MS_NCAND * 65 * sizeof(float), hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(k_chunk_stats, dim3((nchunks + WAVES - 1) / WAVES), dim3(WG), 0, s, d_values, nchunks,
                       d_stats, (const float *)d_cost, d_time_labels);
    MS_HIP(hipGetLastError());
    if (cost_table) MS_HIP(hipStreamSynchronize(s));
    return MS_OK;
}

extern "C" int ms_chunk_features(const uint64_t *d_stats, uint64_t nchunks, float *d_features, uint64_t *d_sizes,
                                 void *stream)
{
    MS_TRY(ms_need_init());
    if (!d_stats) return ms_fail(MS_ERR_INVALID, "null argument");
    if (nchunks == 0) return MS_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_chunk_features, dim3((nchunks + WG - 1) / WG), dim3(WG), 0, s, d_stats, nchunks,
                       d_features, d_sizes);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

extern "C" int ms_candidate_format(uint32_t cand, uint32_t static_bw, ms_format_desc *out)
{
    if (!out || cand >= MS_NCAND) return ms_fail(MS_ERR_INVALID, "candidate %u", cand);
    ms_format_desc f = {};
    switch (cand) {
    case MS_CAND_STATIC:
        if (static_bw < 1 || static_bw > 64) return ms_fail(MS_ERR_INVALID, "static bit width %u", static_bw);
        f.kind = MS_STATIC_VBP; f.bw = static_bw; f.step = 8;
        break;
    case MS_CAND_DVBP:
        f.kind = MS_DYNAMIC_VBP; f.block = 512; f.page = 64; f.step = 8;
        break;
    case MS_CAND_DELTA:
    case MS_CAND_FOR:
        f.kind = cand == MS_CAND_DELTA ? MS_DELTA : MS_FOR; f.block = 1024; f.page = 8; f.step = 8;
        f.inner_kind = MS_DYNAMIC_VBP; f.inner_block = 512; f.inner_page = 64; f.inner_step = 8;
        break;
    default:
        f.kind = MS_RLE;
        break;
    }
    *out = f;
    return MS_OK;
}

struct ms_mlp {
    float *prm = nullptr, *m1 = nullptr, *m2 = nullptr;   // device
    uint64_t steps = 0;                                    // Adam steps taken
};

extern "C" int ms_mlp_create(ms_mlp **out, uint64_t seed)
{
    MS_TRY(ms_need_init());
    if (!out) return ms_fail(MS_ERR_INVALID, "null argument");
    ms_mlp *m = new (std::nothrow) ms_mlp();
    if (!m) return ms_fail(MS_ERR_NOMEM, "host allocation");
    if (hipMalloc((void **)&m->prm, NPARAM * 4) != hipSuccess || hipMalloc((void **)&m->m1, NPARAM * 4) != hipSuccess ||
        hipMalloc((void **)&m->m2, NPARAM * 4) != hipSuccess) {
        (void)hipFree(m->prm); (void)hipFree(m->m1); (void)hipFree(m->m2);
        delete m;
        return ms_fail(MS_ERR_NOMEM, "mlp parameters");
    }
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)), biases 0, from splitmix64
    std::vector<float> h(NPARAM, 0.f);
    auto fill = [&](int off, int fan_in, int fan_out, uint64_t salt) {
        const double a = sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_in * fan_out; i++) {
            uint64_t z = seed + salt + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            h[off + i] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(OW1, IN, HID, 0x1000);
    fill(OW2, HID, HID, 0x2000);
    fill(OW3, HID, OUT, 0x3000);
    if (hipMemcpy(m->prm, h.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, NPARAM * 4) != hipSuccess || hipMemset(m->m2, 0, NPARAM * 4) != hipSuccess) {
        (void)hipFree(m->prm); (void)hipFree(m->m1); (void)hipFree(m->m2);
        delete m;
        return ms_fail(MS_ERR_HIP, "mlp initialisation");
    }
    *out = m;
    return MS_OK;
}

extern "C" int ms_mlp_get_params(const ms_mlp *m, float *host_params)
{
    if (!m || !host_params) return ms_fail(MS_ERR_INVALID, "null argument");
    MS_HIP(hipDeviceSynchronize());
    MS_HIP(hipMemcpy(host_params, m->prm, NPARAM * 4, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_mlp_set_params(ms_mlp *m, const float *host_params)
{
    if (!m || !host_params) return ms_fail(MS_ERR_INVALID, "null argument");
    MS_HIP(hipDeviceSynchronize());
    MS_HIP(hipMemcpy(m->prm, host_params, NPARAM * 4, hipMemcpyHostToDevice));
    MS_HIP(hipMemset(m->m1, 0, NPARAM * 4));
    MS_HIP(hipMemset(m->m2, 0, NPARAM * 4));
    m->steps = 0;
    return MS_OK;
}

// Forward + backward over samples [first, first + count) of `order` (or
// of the identity), then reduction (and Adam when adam != 0).
static int mlp_pass(ms_mlp *m, const float *d_x, const float *d_y, const uint32_t *d_order, uint64_t first,
                    uint64_t count, const float *d_w, float norm, int adam, float lr, float *d_loss,
                    float loss_scale, float *d_grad, float *part, float *lpart, hipStream_t s)
{
    const uint32_t nwg = (uint32_t)((count + SPW - 1) / SPW);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nwg), dim3(WG), 0, s, (const float *)m->prm, d_x, d_y, d_order, first,
                       count, d_w, norm, part, lpart);
    MS_HIP(hipGetLastError());
    float c1 = 1.f, c2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        m->steps++;
        c1 = 1.f - powf(b1, (float)m->steps);
        c2 = 1.f - powf(b2, (float)m->steps);
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 1 + WG - 1) / WG), dim3(WG), 0, s, m->prm, m->m1, m->m2,
                       (const float *)part, nwg, (const float *)lpart, loss_scale, d_loss, adam, lr, b1, b2, 1e-8f,
                       c1, c2, d_grad);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

extern "C" int ms_mlp_loss_grad(ms_mlp *m, const float *d_x, const float *d_y, const float *out_weights,
                                uint64_t n, float *d_grad, float *d_loss, void *stream)
{
    if (!m || !d_x || !d_y || !out_weights || !d_grad || !d_loss || n == 0)
        return ms_fail(MS_ERR_INVALID, "null argument");
    if (n > (1ull << 31)) return ms_fail(MS_ERR_CAPACITY, "too many samples");
    hipStream_t s = (hipStream_t)stream;
    temps sc(s);
    const uint64_t nwg = (n + SPW - 1) / SPW;
    float *part = (float *)sc.get(nwg * NPARAM * 4);
    float *lpart = (float *)sc.get(nwg * 4);
    float *d_w = (float *)sc.get(OUT * 4);
    if (!part || !lpart || !d_w) return ms_fail(MS_ERR_NOMEM, "mlp scratch");
    MS_HIP(hipMemcpyAsync(d_w, out_weights, OUT * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, d_loss, 1u, 0.f);
    MS_TRY(mlp_pass(m, d_x, d_y, nullptr, 0, n, d_w, (float)n, 0, 0.f, d_loss, 1.f / (float)n, d_grad, part, lpart,
                    s));
    MS_HIP(hipStreamSynchronize(s));
    return MS_OK;
}

extern "C" int ms_mlp_train(ms_mlp *m, const float *d_x, const float *d_y, const float *out_weights, uint64_t n,
                            uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *epoch_loss,
                            void *stream)
{
    if (!m || !d_x || !d_y || !out_weights || n == 0 || batch == 0 || batch % SPW)
        return ms_fail(MS_ERR_INVALID, "bad training arguments (batch must be a multiple of %d)", SPW);
    if (n > (1ull << 31)) return ms_fail(MS_ERR_CAPACITY, "too many samples");
    hipStream_t s = (hipStream_t)stream;
    temps sc(s);
    const uint64_t nwg = batch / SPW;
    float *part = (float *)sc.get(nwg * NPARAM * 4);
    float *lpart = (float *)sc.get(nwg * 4);
    float *d_w = (float *)sc.get(OUT * 4);
    float *d_loss = (float *)sc.get((uint64_t)(epochs ? epochs : 1) * 4);
    uint32_t *d_order = (uint32_t *)sc.get(n * 4);
    if (!part || !lpart || !d_w || !d_loss || !d_order) return ms_fail(MS_ERR_NOMEM, "mlp scratch");
    MS_HIP(hipMemcpyAsync(d_w, out_weights, OUT * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, d_loss, epochs, 0.f);
    std::vector<uint32_t> order(n);
    for (uint32_t e = 0; e < epochs; e++) {
        // Fisher-Yates with splitmix64 of (seed, epoch)
        for (uint64_t i = 0; i < n; i++) order[i] = (uint32_t)i;
        for (uint64_t i = n - 1; i > 0; i--) {
            uint64_t z = seed + ((uint64_t)e << 40) + i * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const uint64_t j = z % (i + 1);
            const uint32_t tmp = order[i]; order[i] = order[j]; order[j] = tmp;
        }
        MS_HIP(hipMemcpyAsync(d_order, order.data(), n * 4, hipMemcpyHostToDevice, s));
        for (uint64_t b0 = 0; b0 < n; b0 += batch) {
            const uint64_t cnt = n - b0 < batch ? n - b0 : batch;
            MS_TRY(mlp_pass(m, d_x, d_y, d_order, b0, cnt, d_w, (float)cnt, 1, lr, d_loss + e, 1.f / (float)n,
                            nullptr, part, lpart, s));
        }
        MS_HIP(hipStreamSynchronize(s));   // `order` is rewritten next epoch
    }
    if (epoch_loss && epochs) MS_HIP(hipMemcpy(epoch_loss, d_loss, epochs * 4, hipMemcpyDeviceToHost));
    return MS_OK;
}

extern "C" int ms_mlp_predict(const ms_mlp *m, const float *d_x, uint64_t n, float *d_y, void *stream)
{
    if (!m || !d_x || !d_y) return ms_fail(MS_ERR_INVALID, "null argument");
    if (n == 0) return MS_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mlp_predict, dim3((n + WG - 1) / WG), dim3(WG), 0, s, (const float *)m->prm, d_x, n, d_y);
    MS_HIP(hipGetLastError());
    return MS_OK;
}

extern "C" int ms_mlp_free(ms_mlp *m)
{
    if (!m) return MS_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(m->prm);
    (void)hipFree(m->m1);
    (void)hipFree(m->m2);
    delete m;
    return MS_OK;
}

extern "C" int ms_select_formats(const float *d_pred, const uint64_t *d_sizes, uint64_t n,
                                 const ms_select_params *sp, uint8_t *d_choice, void *stream)
{
    if (!d_pred || !d_sizes || !sp || !d_choice) return ms_fail(MS_ERR_INVALID, "null argument");
    if (n == 0) return MS_OK;
    hipStream_t s = (hipStream_t)stream;
    const float4 norm = make_float4(sp->size_off, sp->size_scale, sp->time_off, sp->time_scale);
    if (sp->rule > 1) return ms_fail(MS_ERR_INVALID, "selection rule %u", sp->rule);
    hipLaunchKernelGGL(k_select_formats, dim3((n + WG - 1) / WG), dim3(WG), 0, s, d_pred, d_sizes, n, sp->size_tol,
                       sp->exact_tol, sp->time_margin, norm, sp->rule, d_choice);
    MS_HIP(hipGetLastError());
    return MS_OK;
}
