"""CPU restatement of the chunk generator, chunk statistics, features and
exact candidate sizes of the learned format selection (numpy, integer
arithmetic except the float features).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the HIP kernels.

The sizes restate the byte layouts of the candidate formats for a column of
exactly 2^16 values (no uncompressed rest, column.h:182-196):
static_vbp.h:60-163, dynamic_vbp.h:83-371 (page of `pbs` meta bytes in front
of up to `pbs` blocks, dynamic_vbp.h:251-255), delta.h:78-490 and
for.h:84-467 (one inner morph_batch call per outer block, for_f reference
pages of 8 values), rle.h:42-297 (16 bytes per run).  They are pinned
against the oracle's own byte images (tests/test_learn_oracle.py).
"""
import numpy as np

CHUNK = 1 << 16
BLOCKS = CHUNK // 512
GEO_LEN = 1024
NCAND = 5
NFEAT = 24
(GEN_UNIFORM, GEN_DVBP, GEN_SORTED, GEN_FOR, GEN_RLE) = range(5)
(ST_SUM_BW, ST_MAX_BW, ST_SUM_DBW, ST_MAX_DBW, ST_SUM_FBW, ST_MAX_FBW, ST_RUNS, ST_NONDEC, ST_ZEROS, ST_MIN,
 ST_MAX, ST_SUM_BW2, ST_DLESS, ST_FLESS, ST_HIST, ST_RESERVED) = range(16)
GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def splitmix(seed, idx):
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M64) + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def geo_table(a):
    """thr[0] = 2^64 - 1, thr[k] = floor(thr[k-1] * (a - 1) / a): P(G > k) ~ (1 - 1/a)^k."""
    t = [M64]
    for _ in range(1, GEO_LEN):
        t.append(t[-1] * (a - 1) // a)
    return np.array(t, dtype=np.uint64)


def geo(thr, r):
    """1 + #{k >= 1 : r < thr[k]} (thr decreasing)."""
    asc = thr[1:][::-1]
    return (1 + (len(asc) - np.searchsorted(asc, r, side="right"))).astype(np.uint64)


def top_bits(r, bits):
    if bits == 0:
        return np.zeros_like(r)
    return r >> np.uint64(64 - bits)


def gen_chunk(kind, a, b, seed, base, thr=None):
    i = np.arange(CHUNK, dtype=np.uint64)
    if kind == GEN_UNIFORM:
        return top_bits(splitmix(seed, i), a)
    if kind == GEN_DVBP:
        blk = np.arange(BLOCKS, dtype=np.uint64)
        bw = np.uint64(a) + splitmix(seed ^ 0xB1, blk) % np.uint64(b - a + 1)
        force = (splitmix(seed ^ 0xB2, blk) % np.uint64(512)).astype(np.int64)
        r = splitmix(seed, i).reshape(BLOCKS, 512)
        x = r >> (np.uint64(64) - bw)[:, None]
        x[np.arange(BLOCKS), force] |= np.uint64(1) << (bw - np.uint64(1))
        return x.reshape(-1)
    if kind == GEN_SORTED:
        g = geo(thr, splitmix(seed, i))
        g[0] = 0
        with np.errstate(over="ignore"):
            return np.uint64(base) + np.cumsum(g, dtype=np.uint64)
    if kind == GEN_FOR:
        with np.errstate(over="ignore"):
            return np.uint64(base) + top_bits(splitmix(seed, i), a)
    # GEN_RLE
    out = np.empty(CHUNK, np.uint64)
    pos, k0 = 0, 0
    while pos < CHUNK:
        k = np.arange(k0, k0 + 4096, dtype=np.uint64)
        lens = geo(thr, splitmix(seed ^ 0x77, k)).astype(np.int64)
        vals = top_bits(splitmix(seed ^ 0x78, k), b)
        rep = np.repeat(vals, lens)
        take = min(len(rep), CHUNK - pos)
        out[pos:pos + take] = rep[:take]
        pos += take
        k0 += 4096
    return out


def eff_bw(v):
    """Bits needed for each v (math.h:108-113), 0 for 0."""
    v = np.asarray(v, dtype=np.uint64).copy()
    n = np.zeros(v.shape, np.int64)
    for s in (32, 16, 8, 4, 2, 1):
        m = (v >> np.uint64(s)) != 0
        n += m * s
        v = np.where(m, v >> np.uint64(s), v)
    return n + (v != 0)


def block_or(x):
    return np.bitwise_or.reduce(x.reshape(-1, 512), axis=1)


def for_refs(x):
    """for.h:121-161 with 8 lanes: signed minimum per lane, unsigned minimum
    over lanes; 2^64 - 1 becomes 2^64 - 2."""
    lanes = x.reshape(-1, 128, 8).view(np.int64).min(axis=1).view(np.uint64)
    r = lanes.min(axis=1)
    r[r == np.uint64(M64)] = np.uint64(M64 - 1)
    return r


def stats(x):
    x = np.asarray(x, dtype=np.uint64)
    assert len(x) == CHUNK
    with np.errstate(over="ignore"):
        bw = eff_bw(block_or(x) | np.uint64(1))
        prev8 = np.concatenate([np.zeros(8, np.uint64), x[:-8]])
        dbw = eff_bw(block_or(x - prev8) | np.uint64(1))
        refs = np.repeat(for_refs(x), 1024)
        fbw = eff_bw(block_or(x - refs) | np.uint64(1))
    s = np.zeros(16, dtype=object)
    s[ST_SUM_BW], s[ST_MAX_BW] = int(bw.sum()), int(bw.max())
    s[ST_SUM_DBW], s[ST_MAX_DBW] = int(dbw.sum()), int(dbw.max())
    s[ST_SUM_FBW], s[ST_MAX_FBW] = int(fbw.sum()), int(fbw.max())
    s[ST_RUNS] = 1 + int(np.count_nonzero(x[1:] != x[:-1]))
    s[ST_NONDEC] = int(np.count_nonzero(x[1:] >= x[:-1]))
    s[ST_ZEROS] = int(np.count_nonzero(x == 0))
    s[ST_MIN], s[ST_MAX] = int(x.min()), int(x.max())
    s[ST_SUM_BW2] = int((bw * bw).sum())
    s[ST_DLESS] = int(np.count_nonzero(dbw < bw))
    s[ST_FLESS] = int(np.count_nonzero(fbw < bw))
    hist = np.bincount((bw - 1) // 8, minlength=8)
    s[ST_HIST] = sum(int(c) << (8 * k) for k, c in enumerate(hist))
    s[ST_RESERVED] = 0
    return np.array([int(v) for v in s], dtype=np.uint64), bw, dbw, fbw


def sizes(st):
    st = [int(v) for v in st]
    return np.array([CHUNK // 8 * st[ST_MAX_BW],
                     (CHUNK // 512 // 64) * 64 + 64 * st[ST_SUM_BW],
                     (CHUNK // 1024) * 64 + 64 * st[ST_SUM_DBW],
                     (CHUNK // 1024 // 8) * 64 + (CHUNK // 1024) * 64 + 64 * st[ST_SUM_FBW],
                     16 * st[ST_RUNS]], dtype=np.uint64)


def _bits(v):
    return 0 if v == 0 else int(v).bit_length()


def features(st):
    st = [int(v) for v in st]
    nb = float(BLOCKS)
    f = np.zeros(NFEAT)
    for k in range(8):
        f[k] = ((st[ST_HIST] >> (8 * k)) & 0xff) / nb
    f[8] = st[ST_SUM_BW] / nb / 64
    f[9] = st[ST_MAX_BW] / 64
    f[10] = st[ST_SUM_DBW] / nb / 64
    f[11] = st[ST_MAX_DBW] / 64
    f[12] = st[ST_SUM_FBW] / nb / 64
    f[13] = st[ST_MAX_FBW] / 64
    f[14] = np.log2(st[ST_RUNS]) / 16
    f[15] = st[ST_RUNS] / CHUNK
    f[16] = st[ST_DLESS] / nb
    f[17] = st[ST_FLESS] / nb
    f[18] = _bits(st[ST_MAX] - st[ST_MIN]) / 64
    f[19] = _bits(st[ST_MIN]) / 64
    f[20] = _bits(st[ST_MAX]) / 64
    f[21] = st[ST_NONDEC] / (CHUNK - 1)
    mean = st[ST_SUM_BW] / nb
    f[22] = np.sqrt(max(st[ST_SUM_BW2] / nb - mean * mean, 0.0)) / 32
    f[23] = st[ST_ZEROS] / CHUNK
    return f


def time_labels(st, bw, dbw, fbw, cost):
    """Decode-time labels from a cost table (NCAND x 65): ns per 512-value
    unit by bit width (static uses the chunk's maximum for every block);
    rle: cost[4][0] + cost[4][1] * runs."""
    cost = np.asarray(cost, dtype=np.float64).reshape(NCAND, 65)
    return np.array([BLOCKS * cost[0, int(st[ST_MAX_BW])], cost[1, bw].sum(), cost[2, dbw].sum(),
                     cost[3, fbw].sum(), cost[4, 0] + cost[4, 1] * float(st[ST_RUNS])])
