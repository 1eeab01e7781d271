"""Independent formulations used to pin the oracle.

Nothing here calls the oracle.  The vertical bit-packed layout is rebuilt
from bit strings (one LSB-first bit string per lane, cut into 64-bit words),
the unmasked packing is re-derived by simulating the operation sequence the
reference's generator emits (include/core/morphing/vbp_routine_gen.py:61-119),
and the operators are written with numpy / collections primitives.
Reference paths are relative to the reference root.
"""
import collections

import numpy as np

M64 = (1 << 64) - 1


# ---------------------------------------------------------------- vbp_l
def lane_bitstrings(values, bw, S):
    lanes = []
    for j in range(S):
        codes = [int(v) for v in values[j::S]]
        lanes.append("".join(format(c, f"0{bw}b")[::-1] for c in codes))
    return lanes


def bitstring_pack(values, bw, S):
    """vbp_l<bw,S> image built from per-lane bit strings (values must fit)."""
    lanes = lane_bitstrings(values, bw, S)
    nwords = len(lanes[0]) // 64
    words = []
    for w in range(nwords):
        for j in range(S):
            words.append(int(lanes[j][64 * w:64 * w + 64][::-1], 2))
    return np.array(words, dtype=np.uint64).view(np.uint8)


def bitstring_unpack(image, n, bw, S):
    """Reads n codes of vbp_l<bw,S> bit by bit."""
    words = np.frombuffer(np.ascontiguousarray(image).tobytes(), dtype=np.uint64)
    per_lane = n // S
    nwords = per_lane * bw // 64
    out = np.zeros(n, dtype=np.uint64)
    for j in range(S):
        bits = "".join(format(int(words[w * S + j]), "064b")[::-1] for w in range(nwords))
        for k in range(per_lane):
            out[k * S + j] = int(bits[k * bw:(k + 1) * bw][::-1], 2)
    return out


def min_cycle_len(bw):
    c = 64
    while not bw & 1:
        bw >>= 1
        c >>= 1
    return c


def emitted_pack(values, bw, S):
    """Simulates the compression routine vbp_routine_gen.py:61-119 emits:
    a register tmp collects unmasked `v << bitpos`; it is stored when full
    and continues with `v >> (64 - old bitpos)` when a code crosses words."""
    values = [int(v) for v in values]
    cyc = min_cycle_len(bw)
    n = len(values)
    out = []
    tmp = [0] * S
    base = 0
    for _ in range(n // (cyc * S)):
        bitpos = 0
        for pos in range(cyc):
            tmp2 = values[base:base + S]
            base += S
            tmp = [(t | (v << bitpos)) & M64 for t, v in zip(tmp, tmp2)]
            bitpos += bw
            if (pos + 1) * bw % 64 == 0:
                out.extend(tmp)
                tmp = [0] * S
                bitpos = 0
            elif (pos * bw) // 64 < ((pos + 1) * bw - 1) // 64:
                out.extend(tmp)
                tmp = [(v >> (bw - bitpos + 64)) & M64 for v in tmp2]
                bitpos -= 64
    return np.array(out, dtype=np.uint64).view(np.uint8)


# ---------------------------------------------------------- dynamic_vbp
def block_bw(block):
    o = 1
    for v in block:
        o |= int(v)
    return o.bit_length()


def dvbp_walk(image, n, B, P, S):
    """Walks pages of one dynamic_vbp_f<B,P,S> morph_batch image: returns the
    decoded values, the meta bytes per page and the bytes consumed."""
    image = np.ascontiguousarray(image, dtype=np.uint8)
    pos, done, vals, metas = 0, 0, [], []
    while done < n:
        meta = image[pos:pos + P].copy()
        metas.append(meta)
        pos += P
        for q in range(P):
            bw = int(meta[q])
            if bw == 0xFF:
                continue
            size = B * bw // 8
            vals.append(bitstring_unpack(image[pos:pos + size], B, bw, S))
            pos += size
            done += B
    return (np.concatenate(vals) if vals else np.zeros(0, np.uint64)), metas, pos


def dvbp_expected(values, B, P, S):
    """dynamic_vbp_f image from the format's definition
    (include/core/morphing/dynamic_vbp.h:188-257)."""
    values = np.asarray(values, dtype=np.uint64)
    out = bytearray()
    nb = len(values) // B
    for p0 in range(0, nb, P):
        blocks = range(p0, min(nb, p0 + P))
        meta = bytearray([0xFF] * P)
        payload = bytearray()
        for q, b in enumerate(blocks):
            blk = values[b * B:(b + 1) * B]
            bw = block_bw(blk)
            meta[q] = bw
            payload += bitstring_pack(blk, bw, S).tobytes()
        out += meta + payload
    return np.frombuffer(bytes(out), dtype=np.uint8)


# ------------------------------------------------------------- cascades
def lane_delta(values, S):
    """d_i = x_i - x_{i-S} mod 2^64 with x_{<0} = 0 (delta.h:124-141)."""
    x = np.asarray(values, dtype=np.uint64)
    prev = np.concatenate([np.zeros(S, np.uint64), x[:-S]]) if len(x) >= S else np.zeros(len(x), np.uint64)
    with np.errstate(over="ignore"):
        return x - prev


def for_reference(block, S):
    """for.h:121-161: lane-wise minimum (signed when S > 1), then the
    unsigned minimum over lanes; 2^64-1 becomes 2^64-2."""
    block = np.asarray(block, dtype=np.uint64)
    if S == 1:
        r = int(block.min())
    else:
        lane_min = block.reshape(-1, S).view(np.int64).min(axis=0).view(np.uint64)
        r = int(lane_min.min())
    return r - 1 if r == M64 else r


def rle_runs(values):
    x = np.asarray(values, dtype=np.uint64)
    if len(x) == 0:
        return np.zeros((0, 2), np.uint64)
    starts = np.concatenate([[0], np.flatnonzero(x[1:] != x[:-1]) + 1])
    ends = np.concatenate([starts[1:], [len(x)]])
    return np.stack([x[starts], (ends - starts).astype(np.uint64)], axis=1)


# ------------------------------------------------------------ operators
CMP = {
    0: lambda x, v: x == v, 1: lambda x, v: x != v, 2: lambda x, v: x < v,
    3: lambda x, v: x <= v, 4: lambda x, v: x > v, 5: lambda x, v: x >= v,
}


def select_positions(x, cmp, val):
    x = np.asarray(x, dtype=np.uint64)
    return np.flatnonzero(CMP[cmp](x, np.uint64(val))).astype(np.uint64)


def masked_sum(x, cmp, val):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(np.where(CMP[cmp](x, np.uint64(val)), x, np.uint64(0)), dtype=np.uint64))


def multiset_intersect(a, b):
    cb = collections.Counter(int(v) for v in b)
    out = []
    for v, c in sorted(collections.Counter(int(v) for v in a).items()):
        out += [v] * min(c, cb.get(v, 0))
    return np.array(out, dtype=np.uint64)


def multiset_union(a, b):
    ca = collections.Counter(int(v) for v in a)
    cb = collections.Counter(int(v) for v in b)
    out = []
    for v in sorted(set(ca) | set(cb)):
        out += [v] * max(ca.get(v, 0), cb.get(v, 0))
    return np.array(out, dtype=np.uint64)


def calc(op, a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    with np.errstate(over="ignore", divide="ignore"):
        nz = np.where(b == 0, np.uint64(1), b)
        return {
            0: lambda: a + b, 1: lambda: a - b, 2: lambda: a * b,
            3: lambda: np.where(b == 0, np.uint64(0), a // nz),
            4: lambda: np.where(b == 0, np.uint64(0), a % nz),
            5: lambda: np.minimum(a, b), 6: lambda: np.maximum(a, b),
            7: lambda: a & b, 8: lambda: a | b, 9: lambda: a ^ b,
        }[op]()


def align64(x):
    return (x + 63) // 64 * 64
