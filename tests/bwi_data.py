"""Synthetic columns for blockwise_individual_f: every block drawn from one
of several value distributions (small uniform, sorted with small gaps,
runs, a large base plus small offsets, wide values, zeros), so that every
default format wins some block and the selectors have real choices."""
import numpy as np

from engine_amd import datagen


def block_values(seed, kind, B):
    u = datagen.uniform_bits
    if kind == 0:
        return u(seed, B, 1 + int(u(seed + 1, 1, 5)[0]) % 30)
    if kind == 1:
        return np.cumsum(1 + u(seed, B, 3)).astype(np.uint64)
    if kind == 2:
        runs = u(seed, B // 64 + 1, 20)
        return np.repeat(runs, 64)[:B].astype(np.uint64)
    if kind == 3:
        return (np.uint64(1 << 40) + u(seed, 1, 40)[0] + u(seed + 1, B, 12)).astype(np.uint64)
    if kind == 4:
        return u(seed, B, 63)
    return np.zeros(B, np.uint64)


def mixture(seed, nblocks, B, rest=0, kinds=6):
    parts = [block_values(seed + 17 * b, int(datagen.uniform_bits(seed + b, 1, 16)[0]) % kinds, B)
             for b in range(nblocks)]
    parts.append(datagen.uniform_bits(seed + 99, rest, 20))
    return np.ascontiguousarray(np.concatenate(parts).astype(np.uint64))


def fitting_codes(x, B, seed, allowed):
    """A code per block from `allowed`, static_vbp codes raised to the
    block's maximum bit width (narrower ones would not hold its values)."""
    nb = len(x) // B
    pick = datagen.uniform_bits(seed, nb, 16) % np.uint64(len(allowed))
    codes = np.array([allowed[int(i)] for i in pick], np.uint8)
    for b in range(nb):
        if 1 <= codes[b] <= 64:
            mbw = max(1, int(np.bitwise_or.reduce(x[b * B:(b + 1) * B])).bit_length())
            codes[b] = max(int(codes[b]), mbw)
    return codes
