"""Seeded synthetic columns for the workloads (integer arithmetic, numpy).

All randomness comes from splitmix64 of a counter: value r_i of stream
`seed` is splitmix64(seed + (i + 1) * 0x9E3779B97F4A7C15).  The base seed is
0x4D6F7270 plus the workload index; every generated column can be hashed
with `column_hash` so that two runs can show they used the same input.

Workloads (uint64 columns):
  select_sum_12bit   2^24 values uniform in [0, 2^12)
  dvbp_random_bw     per 512-value block a bit width b ~ U{1..32}, values
                     uniform in [0, 2^b) with one value of the block forced to
                     have bit b-1 set
  sorted_geometric   x_0 = 0, x_i = x_{i-1} + g, g ~ geometric(p = 0.1) >= 1
  for_base_offset    x = B + (r >> 48), B ~ U[2^40, 2^41) drawn once
  rle_runs           run lengths ~ geometric(mean 64), values U[0, 2^20)
  mixed_columns      columns of log-uniform length, each in one of five
                     formats with the matching value distribution above
"""
import hashlib

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 0x4D6F7270


def splitmix64(seed, n, offset=0):
    """n values of the splitmix64 stream `seed`, starting at index `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def column_hash(x):
    return hashlib.blake2b(np.ascontiguousarray(x, dtype=np.uint64).tobytes(), digest_size=8).hexdigest()


def uniform_bits(seed, n, bits):
    """Uniform in [0, 2^bits)."""
    return splitmix64(seed, n) >> np.uint64(64 - bits) if bits > 0 else np.zeros(n, np.uint64)


def dvbp_random_bw(seed, n, block=512, lo=1, hi=32):
    assert n % block == 0
    nb = n // block
    b = (splitmix64(seed ^ 0xB1, nb) % np.uint64(hi - lo + 1) + np.uint64(lo)).astype(np.uint64)
    r = splitmix64(seed, n).reshape(nb, block)
    shift = (np.uint64(64) - b)[:, None]
    x = r >> shift
    # force bit b-1 in one value per block, so the block's bit width is exactly b
    k = (splitmix64(seed ^ 0xB2, nb) % np.uint64(block)).astype(np.int64)
    rows = np.arange(nb)
    x[rows, k] |= np.uint64(1) << (b - np.uint64(1))
    return x.reshape(-1)


def geometric(seed, n, p_num, p_den):
    """Geometric on {1, 2, ...} with success probability p_num/p_den, by
    inversion of a 53-bit uniform (deterministic for a given numpy)."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u = np.maximum(u, 1e-300)
    p = p_num / p_den
    g = np.floor(np.log(u) / np.log1p(-p)) + 1.0
    return np.minimum(g, 1e9).astype(np.uint64)


def sorted_geometric(seed, n, p_num=1, p_den=10):
    g = geometric(seed, n, p_num, p_den)
    if n == 0:
        return g
    g[0] = 0
    return np.cumsum(g, dtype=np.uint64)


def for_base_offset(seed, n, offset_bits=16):
    base = (1 << 40) + int(splitmix64(seed ^ 0xF0, 1)[0] >> np.uint64(24))
    return np.uint64(base) + (splitmix64(seed, n) >> np.uint64(64 - offset_bits))


def rle_runs(seed, n, mean_len=64, value_bits=20):
    out = np.empty(n, np.uint64)
    if n == 0:
        return out
    pos, chunk, off = 0, max(1024, n // mean_len + 16), 0
    while pos < n:
        lens = geometric(seed ^ 0x77 ^ (off * 0x9E3779B1), chunk, 1, mean_len)
        vals = splitmix64(seed ^ 0x78, chunk, off) >> np.uint64(64 - value_bits)
        off += chunk
        rep = np.repeat(vals, lens.astype(np.int64))
        take = min(len(rep), n - pos)
        out[pos:pos + take] = rep[:take]
        pos += take
    return out


def zipf_flags(seed, n, k=8, s=1.1):
    """Values 0..k-1 with P(v) proportional to 1/(v+1)^s."""
    w = 1.0 / np.arange(1, k + 1) ** s
    cdf = np.cumsum(w / w.sum())
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.searchsorted(cdf, u, side="right").clip(0, k - 1).astype(np.uint64)


MIXED_KINDS = ("static_vbp", "dynamic_vbp", "delta", "for", "rle")


def mixed_columns(seed, ncols, lo_exp=3.0, hi_exp=6.0):
    """Yields (kind, format, values, select value) for `ncols` columns:
    length floor(10^(lo_exp + (hi_exp - lo_exp) u)), format uniform over
    static_vbp<b, 8> (b ~ U{1..32}), dynamic_vbp<512,2048,8>,
    delta<1024,8,dynamic_vbp<512,64,8>>, for<1024,8,dynamic_vbp<512,64,8>>
    and rle, values drawn from that format's generator; the select value
    (for x < value) passes about half of the values."""
    from .formats import (default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, dynamic_vbp_f, rle_f,
                          static_vbp_f)
    r = splitmix64(seed ^ 0xC4, 3 * ncols).reshape(ncols, 3)
    for c in range(ncols):
        u = float(r[c, 0] >> np.uint64(11)) / 9007199254740992.0
        n = int(10.0 ** (lo_exp + (hi_exp - lo_exp) * u))
        kind = MIXED_KINDS[int(r[c, 1] % np.uint64(5))]
        s = seed + 0x1000 * (c + 1)
        if kind == "static_vbp":
            b = int(r[c, 2] % np.uint64(32)) + 1
            yield kind, static_vbp_f(b, 8), uniform_bits(s, n, b), 1 << (b - 1)
        elif kind == "dynamic_vbp":
            nb = n // 512 * 512
            x = np.concatenate([dvbp_random_bw(s, nb), uniform_bits(s + 1, n - nb, 16)])
            yield kind, dynamic_vbp_f(512, 2048, 8), x, 1 << 16
        elif kind == "delta":
            x = sorted_geometric(s, n)
            yield kind, default_delta_dynamic_vbp_f(8), x, int(x[n // 2]) if n else 1
        elif kind == "for":
            x = for_base_offset(s, n)
            yield kind, default_for_dynamic_vbp_f(8), x, (int(x.min()) if n else 0) + (1 << 15)
        else:
            yield kind, rle_f(), rle_runs(s, n), 1 << 19
