"""Pins the CPU oracle against the reference's known-answer tests, closed
forms, bit-level re-derivations and brute force (no GPU needed)."""
import json
import os

import numpy as np
import pytest

from engine_amd import datagen
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_GE, MS_CMP_GT, MS_CMP_LE, MS_CMP_LT, MS_CMP_NE,
                                 MS_OP_ADD, default_delta_dynamic_vbp_f, default_dynamic_vbp_f,
                                 default_for_dynamic_vbp_f, delta_f, dynamic_vbp_f, for_f, rle_f,
                                 static_vbp_f, uncompr_f)
from tests import refmodel as rm

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "reference_kat.json")))


def rng_vals(seed, n, bits):
    return datagen.uniform_bits(seed, n, bits)


# ------------------------------------------------------------------ KATs
def test_kat_select(oracle):
    k = KAT["select_less"]
    out = oracle.select(MS_CMP_LT, k["val"], uncompr_f(), oracle.from_values(k["in"]), 1)
    assert list(out.values()) == k["out"]
    assert out.meta() == (2, 16, 0)


def test_kat_project(oracle):
    k = KAT["project"]
    out = oracle.project(uncompr_f(), oracle.from_values(k["data"]), oracle.from_values(k["pos"]), 1)
    assert list(out.values()) == k["out"]


def test_kat_agg_sum(oracle):
    k = KAT["agg_sum"]
    out = oracle.agg_sum(oracle.from_values(k["in"]))
    assert list(out.values()) == k["out"]
    k = KAT["agg_sum_1_to_100"]
    lo, hi = k["in_range"]
    assert list(oracle.agg_sum(oracle.from_values(range(lo, hi + 1))).values()) == k["out"]


def test_kat_intersect_merge_calc(oracle):
    k = KAT["intersect_sorted"]
    out = oracle.intersect_sorted(uncompr_f(), oracle.from_values(k["l"]), oracle.from_values(k["r"]))
    assert list(out.values()) == k["out"]
    k = KAT["merge_sorted"]
    out = oracle.merge_sorted(uncompr_f(), oracle.from_values(k["l"]), oracle.from_values(k["r"]))
    assert list(out.values()) == k["out"]
    k = KAT["calc_binary_plus"]
    out = oracle.calc_binary(MS_OP_ADD, uncompr_f(), oracle.from_values(k["l"]), oracle.from_values(k["r"]))
    assert list(out.values()) == k["out"]


# --------------------------------------------------------------- vbp_l
@pytest.mark.parametrize("S", [1, 2, 4, 8])
@pytest.mark.parametrize("bw", [1, 2, 3, 7, 12, 16, 31, 32, 33, 47, 63, 64])
def test_static_vbp_matches_bit_strings(oracle, bw, S):
    n = 64 * S * 2
    v = rng_vals(bw * 100 + S, n, bw)
    img = oracle.morph_batch_compress(static_vbp_f(bw, S), v)
    assert img.tobytes() == rm.bitstring_pack(v, bw, S).tobytes()
    assert len(img) == n * bw // 8
    back, used = oracle.morph_batch_decompress(static_vbp_f(bw, S), img, n)
    assert used == len(img)
    assert np.array_equal(back, v)
    assert np.array_equal(rm.bitstring_unpack(img, n, bw, S), v)


@pytest.mark.parametrize("bw", [3, 5, 12, 20, 40])
def test_static_vbp_unmasked_pack_matches_emitted_routine(oracle, bw):
    """Values wider than bw corrupt neighbours exactly as the generated
    routine's shift/or sequence does (vbp_routine_gen.py:97-112)."""
    S = 4
    n = 64 * S
    v = rng_vals(7 + bw, n, min(64, bw + 6))
    img = oracle.morph_batch_compress(static_vbp_f(bw, S), v)
    assert img.tobytes() == rm.emitted_pack(v, bw, S).tobytes()


# ---------------------------------------------------------- dynamic_vbp
@pytest.mark.parametrize("B,P,S", [(512, 64, 8), (512, 2048, 8), (256, 32, 4), (64, 8, 1), (128, 16, 2)])
@pytest.mark.parametrize("nblocks", [1, 3, 8, 70])
def test_dynamic_vbp_closed_form(oracle, B, P, S, nblocks):
    n = B * nblocks
    v = datagen.dvbp_random_bw(nblocks * 13 + S, n, block=B, lo=1, hi=40)
    v[:B] = 0                                       # an all-zero block packs with bw 1
    f = dynamic_vbp_f(B, P, S)
    img = oracle.morph_batch_compress(f, v)
    assert img.tobytes() == rm.dvbp_expected(v, B, P, S).tobytes()
    vals, metas, used = rm.dvbp_walk(img, n, B, P, S)
    assert used == len(img)
    assert np.array_equal(vals, v)
    assert metas[0][0] == 1
    back, used2 = oracle.morph_batch_decompress(f, img, n)
    assert used2 == len(img) and np.array_equal(back, v)


# ---------------------------------------------------------------- delta
@pytest.mark.parametrize("sorted_", [True, False])
@pytest.mark.parametrize("n", [1024, 3072, 10240])
def test_delta_closed_form(oracle, sorted_, n):
    f = default_delta_dynamic_vbp_f(8)
    v = datagen.sorted_geometric(n, n) if sorted_ else rng_vals(n, n, 30)
    img = oracle.morph_batch_compress(f, v)
    d = rm.lane_delta(v, 8)
    expect = b"".join(rm.dvbp_expected(d[i:i + 1024], 512, 64, 8).tobytes() for i in range(0, n, 1024))
    assert img.tobytes() == expect
    back, used = oracle.morph_batch_decompress(f, img, n)
    assert used == len(img) and np.array_equal(back, v)


# ------------------------------------------------------------------ for
@pytest.mark.parametrize("nblocks", [1, 7, 8, 9, 17])
def test_for_closed_form(oracle, nblocks):
    f = default_for_dynamic_vbp_f(8)
    L, P = 1024, 8
    n = L * nblocks
    v = datagen.for_base_offset(nblocks, n)
    v[5] = np.uint64(2**63 + 5)                     # exercises the signed lane minimum
    img = oracle.morph_batch_compress(f, v)
    pos = 0
    for p0 in range(0, nblocks, P):
        nb = min(P, nblocks - p0)
        refs = np.frombuffer(img[pos:pos + 8 * P].tobytes(), np.uint64)
        pos += 8 * P
        if nb < P:
            assert refs[nb] == np.uint64(2**64 - 1)
        for q in range(nb):
            blk = v[(p0 + q) * L:(p0 + q + 1) * L]
            ref = rm.for_reference(blk, 8)
            assert int(refs[q]) == ref
            with np.errstate(over="ignore"):
                inner = rm.dvbp_expected(blk - np.uint64(ref), 512, 64, 8)
            assert img[pos:pos + len(inner)].tobytes() == inner.tobytes()
            pos += len(inner)
    assert pos == len(img)
    back, used = oracle.morph_batch_decompress(f, img, n)
    assert used == len(img) and np.array_equal(back, v)


def test_for_reference_value_special_case(oracle):
    f = for_f(1024, 8, dynamic_vbp_f(512, 64, 8))
    v = np.full(1024, 2**64 - 1, np.uint64)
    img = oracle.morph_batch_compress(f, v)
    assert int(np.frombuffer(img[:8].tobytes(), np.uint64)[0]) == 2**64 - 2
    assert np.array_equal(oracle.morph_batch_decompress(f, img, 1024)[0], v)


# ------------------------------------------------------------------ rle
@pytest.mark.parametrize("n", [0, 1, 2, 100, 5000])
def test_rle_maximal_runs(oracle, n):
    v = datagen.rle_runs(n + 3, n, mean_len=5) if n else np.zeros(0, np.uint64)
    col = oracle.morph(rle_f(), v)
    runs = rm.rle_runs(v)
    assert col.image().tobytes() == runs.reshape(-1).tobytes()
    assert col.meta() == (n, 16 * len(runs), 16 * len(runs))
    assert np.array_equal(col.values(), v)


# ----------------------------------------------------- column layout A.6
def test_gap_sizes_from_format_definitions(oracle):
    # 2 * F::get_size_max_byte(F::m_BlockSize) for the config formats
    assert oracle.rest_offset(dynamic_vbp_f(512, 2048, 8), 0) == 12288
    assert oracle.rest_offset(static_vbp_f(12, 1), 0) == 192
    assert oracle.rest_offset(default_delta_dynamic_vbp_f(8), 0) == 16640
    assert oracle.rest_offset(default_for_dynamic_vbp_f(8), 0) == 16768
    assert oracle.rest_offset(uncompr_f(), 5) == 64


@pytest.mark.parametrize("f", [dynamic_vbp_f(512, 2048, 8), static_vbp_f(12, 1), static_vbp_f(5, 8),
                               default_delta_dynamic_vbp_f(8), default_for_dynamic_vbp_f(8)],
                         ids=repr)
@pytest.mark.parametrize("n", [0, 7, 64, 1000, 4096 + 77])
def test_column_morph_layout(oracle, f, n):
    bits = f.bw if f.bw else 20
    v = rng_vals(n + 1, n, bits)
    col = oracle.morph(f, v)
    bs = f.block_size()
    nc = n // bs * bs
    main = oracle.morph_batch_compress(f, v[:nc]) if nc else np.zeros(0, np.uint8)
    assert col.size_compr == len(main)
    img = col.image()
    assert img[:len(main)].tobytes() == main.tobytes()
    if n - nc:
        r = rm.align64(len(main) + oracle.rest_offset(f, 0))
        assert col.size_used == r + 8 * (n - nc)
        assert np.array_equal(np.frombuffer(img[r:].tobytes(), np.uint64), v[nc:])
        assert col.count_uncompr() == n - nc
    else:
        assert col.size_used == col.size_compr
    assert np.array_equal(col.values(), v)


# ------------------------------------------------ write-iterator A.7
@pytest.mark.parametrize("f", [dynamic_vbp_f(512, 2048, 8), default_dynamic_vbp_f(8),
                               default_delta_dynamic_vbp_f(8), default_for_dynamic_vbp_f(8),
                               static_vbp_f(20, 8)], ids=repr)
@pytest.mark.parametrize("m", [0, 100, 2048, 5000, 6144, 6144 + 1024])
def test_write_iterator_chunking(oracle, f, m):
    v = np.sort(rng_vals(m + 5, m, 20))
    col, app = oracle.wit_emit(f, v)
    K = 2048
    bs = f.block_size()
    parts = [v[i:i + K] for i in range(0, m // K * K, K)]
    rem = m % K
    nc = rem // bs * bs
    if nc:
        parts.append(v[m // K * K:m // K * K + nc])
    if f.kind == 3:   # delta carries x_{i-step} across flushes: one stream
        main = oracle.morph_batch_compress(f, v[:m // K * K + nc]) if parts else np.zeros(0, np.uint8)
    else:
        main = np.concatenate([oracle.morph_batch_compress(f, p) for p in parts]) if parts else np.zeros(0, np.uint8)
    assert col.size_compr == len(main)
    img = col.image()
    assert img[:len(main)].tobytes() == main.tobytes()
    r = rm.align64(len(main) + oracle.rest_offset(f, 0))
    if rem - nc:
        assert col.size_used == r + 8 * (rem - nc) == app
    else:
        assert col.size_used == len(main) and app == r
    assert np.array_equal(col.values(), v)


@pytest.mark.parametrize("f,page_bytes", [(for_f(512, 1024, dynamic_vbp_f(512, 64, 8)), 8 * 1024),
                                          (for_f(512, 8, dynamic_vbp_f(512, 64, 8)), 8 * 8)], ids=repr)
def test_write_iterator_capacity_per_flush(oracle, f, page_bytes):
    """Every 2048-value flush starts a fresh for_f reference page, so the
    output of m full-width values takes m / 2048 pages: far more than one
    page per 1024 * 512 values that a column-level size bound counts.  The
    image must fit its allocation and have exactly that closed-form size."""
    m = 1 << 19
    v = rng_vals(19, m, 64)
    v[::512] |= np.uint64(1 << 63)                  # every FOR block spans the full 64 bits
    col, app = oracle.wit_emit(f, v)
    flushes = m // 2048
    per_flush = page_bytes + 4 * (64 + 512 * 64 // 8)   # 4 FOR blocks: inner page meta + bw-64 payload
    assert col.size_compr == col.size_used == flushes * per_flush
    assert col.size_used <= col.size_alloc
    assert np.array_equal(col.values(), v)


# ------------------------------------------------------------ operators
CMPS = [MS_CMP_EQ, MS_CMP_NE, MS_CMP_LT, MS_CMP_LE, MS_CMP_GT, MS_CMP_GE]


@pytest.mark.parametrize("lanes", [1, 8])
@pytest.mark.parametrize("in_f", [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8),
                                  default_for_dynamic_vbp_f(8)], ids=repr)
@pytest.mark.parametrize("out_f", [uncompr_f(), default_delta_dynamic_vbp_f(8), default_dynamic_vbp_f(8)],
                         ids=repr)
def test_select_brute_force(oracle, lanes, in_f, out_f):
    n = 3 * 4096 + 13
    v = rng_vals(17, n, 12)
    col = oracle.morph(in_f, v)
    for cmp in CMPS:
        out = oracle.select(cmp, 2048, out_f, col, lanes)
        expect = rm.select_positions(v, cmp, 2048)
        assert np.array_equal(out.values(), expect)
        assert oracle.select_sum(cmp, 2048, col) == rm.masked_sum(v, cmp, 2048)
        assert oracle.select_sum_stream(cmp, 2048, col, 3) == rm.masked_sum(v, cmp, 2048)


def test_select_scalar_remainder_layout(oracle):
    """select_compr.h:209-250: with 8 lanes, positions from the input rest's
    last (rest mod 8) values go raw behind the write-iterator's output."""
    f = dynamic_vbp_f(512, 2048, 8)
    n = 512 * 5 + 21                  # rest of 21 values: 16 vectorised, 5 scalar
    v = rng_vals(3, n, 12)
    col = oracle.morph(f, v)
    out_f = default_delta_dynamic_vbp_f(8)
    out = oracle.select(MS_CMP_LT, 4096, out_f, col, 8)         # everything matches
    wit, app = oracle.wit_emit(out_f, np.arange(n - 5, dtype=np.uint64))
    img = out.image()
    assert out.size_compr == wit.size_compr
    assert img[:wit.size_used].tobytes() == wit.image().tobytes()
    assert np.array_equal(np.frombuffer(img[app:].tobytes(), np.uint64), np.arange(n - 5, n, dtype=np.uint64))
    assert out.size_used == app + 40
    # no match among the scalar remainder: sizeUsed ends where the rest would start
    v2 = v.copy()
    v2[-5:] = 4095
    out2 = oracle.select(MS_CMP_LT, 4095, out_f, oracle.morph(f, v2), 8)
    pos2 = rm.select_positions(v2, MS_CMP_LT, 4095)
    wit2, app2 = oracle.wit_emit(out_f, pos2)
    assert out2.size_used == app2 and np.array_equal(out2.values(), pos2)


@pytest.mark.parametrize("data_f", [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8)], ids=repr)
@pytest.mark.parametrize("pos_f", [uncompr_f(), default_delta_dynamic_vbp_f(8)], ids=repr)
def test_project_brute_force(oracle, data_f, pos_f):
    n = 2 * 4096 + 100
    v = rng_vals(5, n, 12)
    p = np.flatnonzero(v < 1500).astype(np.uint64)
    out = oracle.project(uncompr_f(), oracle.morph(data_f, v), oracle.morph(pos_f, p), 8)
    assert np.array_equal(out.values(), v[p])


def test_set_operators_multiset(oracle):
    a = np.sort(rng_vals(1, 3000, 9))
    b = np.sort(rng_vals(2, 2500, 9))
    for out_f in (uncompr_f(), default_delta_dynamic_vbp_f(8)):
        i = oracle.intersect_sorted(out_f, oracle.morph(default_delta_dynamic_vbp_f(8), a), oracle.from_values(b))
        assert np.array_equal(i.values(), rm.multiset_intersect(a, b))
        m = oracle.merge_sorted(out_f, oracle.from_values(a), oracle.from_values(b))
        assert np.array_equal(m.values(), rm.multiset_union(a, b))


def test_calc_all_ops(oracle):
    a = rng_vals(11, 1000, 64)
    b = rng_vals(12, 1000, 64)
    b[::7] = 0
    for op in range(10):
        out = oracle.calc_binary(op, uncompr_f(), oracle.from_values(a), oracle.from_values(b))
        assert np.array_equal(out.values(), rm.calc(op, a, b)), op
    assert oracle.calc_binary(0, uncompr_f(), oracle.from_values(a), oracle.from_values(b[:10])) == 4


# ------------------------------------------- prepare_for_random_access
PRA_FORMATS = [static_vbp_f(12, 1), static_vbp_f(7, 8), dynamic_vbp_f(512, 2048, 8), default_dynamic_vbp_f(8),
               default_delta_dynamic_vbp_f(8), default_for_dynamic_vbp_f(8), for_f(1024, 8, uncompr_f())]


@pytest.mark.parametrize("f", PRA_FORMATS, ids=repr)
@pytest.mark.parametrize("n", [0, 5, 512, 3000, 4096 + 77])
def test_prepare_for_random_access(oracle, f, n):
    """column.h:241-268 by its definition: the zero-padded rest, compressed
    by an independent morph_batch call, sits at sizeCompr; the main part,
    the raw rest and the meta data are unchanged."""
    bits = f.bw if f.kind == 1 else 20
    v = np.sort(rng_vals(n + 3, n, bits)) if f.kind == 3 else rng_vals(n + 3, n, bits)
    col = oracle.morph(f, v)
    before, meta = col.image(), col.meta()
    bs = f.block_size()
    rest = n % bs
    acted = col.prepare_for_random_access()
    assert acted == (rest > 0)
    assert col.meta() == meta
    after = col.image()
    if not rest:
        assert after.tobytes() == before.tobytes()
        return
    ru = -(-rest // bs) * bs
    padded = np.concatenate([v[n - rest:], np.zeros(ru - rest, np.uint64)])
    want = oracle.morph_batch_compress(f, padded)
    c = col.size_compr
    assert after[:c].tobytes() == before[:c].tobytes()
    assert after[c:c + len(want)].tobytes() == want.tobytes()
    r = oracle.rest_offset(f, c)
    assert after[r:].tobytes() == before[r:].tobytes()
    assert c + len(want) <= r
    # the padding behind the rest is zero, and the compressed rest decodes
    # to the padded rest
    raw = col.raw(r + 8 * ru)
    assert not raw[r + 8 * rest:].any()
    dec, used = oracle.morph_batch_decompress(f, after[c:c + len(want)], ru)
    assert used == len(want) and np.array_equal(dec, padded)
    assert np.array_equal(col.values(), v)


def test_prepare_for_random_access_uncompr_and_rle(oracle):
    assert not oracle.from_values(np.arange(10, dtype=np.uint64)).prepare_for_random_access()
    with pytest.raises(RuntimeError):
        oracle.morph(rle_f(), np.arange(10, dtype=np.uint64)).prepare_for_random_access()


@pytest.mark.parametrize("lanes", [1, 8])
def test_between_brute_force(oracle, lanes):
    """between_compr.h: positions with both comparisons, laid out as select
    lays them out (the two equal a select whose predicate is the same)."""
    x = rng_vals(7, 5000 + 3, 12)
    col = oracle.morph(dynamic_vbp_f(512, 2048, 8), x)
    for lo_cmp, lo, hi_cmp, hi in [(MS_CMP_GE, 100, MS_CMP_LE, 900), (MS_CMP_GT, 100, MS_CMP_LT, 900),
                                   (MS_CMP_GE, 900, MS_CMP_LE, 100), (MS_CMP_EQ, 7, MS_CMP_LE, 4095)]:
        o = oracle.between(lo_cmp, lo, hi_cmp, hi, uncompr_f(), col, lanes)
        want = np.flatnonzero(rm.CMP[lo_cmp](x, np.uint64(lo)) & rm.CMP[hi_cmp](x, np.uint64(hi)))
        assert np.array_equal(o.values(), want.astype(np.uint64))
    # a one-sided between is a select
    a = oracle.between(MS_CMP_GE, 0, MS_CMP_LT, 2048, default_delta_dynamic_vbp_f(8), col, lanes)
    b = oracle.select(MS_CMP_LT, 2048, default_delta_dynamic_vbp_f(8), col, lanes)
    assert a.image().tobytes() == b.image().tobytes() and a.meta() == b.meta()


# ------------------------------------------------------------ group, join
def test_kat_group_join(oracle):
    k = KAT["group_unary"]
    g, e = oracle.group(oracle.from_values(k["data"]))
    assert g.values().tolist() == k["gr"] and e.values().tolist() == k["ext"]
    k = KAT["group_binary"]
    g, e = oracle.group(oracle.from_values(k["data"]), oracle.from_values(k["in_gr"]))
    assert g.values().tolist() == k["gr"] and e.values().tolist() == k["ext"]
    k = KAT["join"]
    a, b = oracle.join(oracle.from_values(k["l"]), oracle.from_values(k["r"]))
    assert a.values().tolist() == k["out_l"] and b.values().tolist() == k["out_r"]
    k = KAT["semi_join"]
    assert oracle.semi_join(oracle.from_values(k["l"]), oracle.from_values(k["r"])).values().tolist() == k["out"]


@pytest.mark.parametrize("n", [0, 1, 1000, 20_000])
def test_group_brute_force(oracle, n):
    x = rng_vals(n + 11, n, 6)
    y = rng_vals(n + 12, n, 2)
    for gin in (None, y):
        keys = list(zip(gin.tolist(), x.tolist())) if gin is not None else x.tolist()
        ids, ext = {}, []
        want = []
        for i, kk in enumerate(keys):
            if kk not in ids:
                ids[kk] = len(ids)
                ext.append(i)
            want.append(ids[kk])
        g, e = oracle.group(oracle.from_values(x), oracle.from_values(gin) if gin is not None else None)
        assert g.values().tolist() == want and e.values().tolist() == ext


@pytest.mark.parametrize("na,nb", [(0, 5), (7, 0), (300, 200), (200, 300), (256, 256)])
def test_join_brute_force(oracle, na, nb):
    a = rng_vals(na + 1, na, 5)
    b = rng_vals(nb + 2, nb, 5)
    l, r = oracle.join(oracle.from_values(a), oracle.from_values(b))
    if na >= nb:
        want = [(i, j) for i in range(na) for j in range(nb) if a[i] == b[j]]
    else:   # the outer loop over the longer right input
        want = [(i, j) for j in range(nb) for i in range(na) if a[i] == b[j]]
    assert list(zip(l.values().tolist(), r.values().tolist())) == want
    s = oracle.semi_join(oracle.from_values(a), oracle.from_values(b))
    assert s.values().tolist() == [i for i in range(na) if a[i] in set(b.tolist())]


# ---------------------------------- vbp_padding, k_wise_ns, group_simple
from engine_amd.formats import group_simple_f, k_wise_ns_f, static_vbp_padding_f  # noqa: E402


@pytest.mark.parametrize("bw,S", [(1, 8), (3, 8), (7, 4), (12, 1), (21, 8), (33, 2), (64, 8)])
def test_vbp_padding_closed_form(oracle, bw, S):
    """vbp_padding.h:97-140: lane j's word of a block is the OR of its
    64 / bw codes shifted by k bw; size = blocks x S words."""
    f = static_vbp_padding_f(bw, S)
    per = 64 // bw
    bs = per * S
    v = rng_vals(bw * 7 + S, 5 * bs, bw)
    img = oracle.morph_batch_compress(f, v)
    want = np.zeros(5 * S, np.uint64)
    for b in range(5):
        for j in range(S):
            w = 0
            for k in range(per):
                w |= int(v[b * bs + k * S + j]) << (k * bw)
            want[b * S + j] = w
    assert img.tobytes() == want.tobytes()
    dec, used = oracle.morph_batch_decompress(f, img, 5 * bs)
    assert used == len(img) and np.array_equal(dec, v)


def test_k_wise_ns_closed_form(oracle):
    """k_wise_ns.h:178-212 with the shuffle tables of :49-115: per pair a
    descriptor (leading zero bytes of x1 << 3 | of x0) and the low bytes."""
    v = np.array([0, 1, 255, 256, 2**63, 2**40 + 5, 7, 2**64 - 1], dtype=np.uint64)
    img = oracle.morph_batch_compress(k_wise_ns_f(2), v)
    want = bytearray()
    for a, b in zip(v[0::2].tolist(), v[1::2].tolist()):
        za, zb = [(64 - (x | 1).bit_length()) // 8 for x in (a, b)]
        want.append((zb << 3) | za)
        want += a.to_bytes(8, "little")[:8 - za] + b.to_bytes(8, "little")[:8 - zb]
    assert img.tobytes() == bytes(want)
    v = rng_vals(5, 4000, 37) >> (rng_vals(6, 4000, 6) % np.uint64(37))
    img = oracle.morph_batch_compress(k_wise_ns_f(2), v)
    dec, used = oracle.morph_batch_decompress(k_wise_ns_f(2), img, 4000)
    assert used == len(img) and np.array_equal(dec, v)


def _gs_reference_model(v, E, base=0):
    """group_simple.h:246-365 written out independently in Python: the ring
    buffer of pseudo group maxima, the greedy mode choice, 4-bit selectors,
    aligned data area, incomplete block."""
    nums = [64, 32, 21, 16, 12, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1]
    bits = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 21, 32, 64]
    G = len(v) // E
    gmax = [0] * G
    for g in range(G):
        for j in range(E):
            gmax[g] |= int(v[g * E + j])
    sels, p = [], 0
    while p < G:
        rb = gmax[p:p + 64]
        for i in range(15):
            n, mask = nums[i], (1 << bits[i]) - 1
            maxpos = min(n, len(rb))
            off = 0
            while off < maxpos and rb[off] <= mask:
                off += 1
            if off == maxpos:
                break
        if off != n:
            break
        sels.append(i)
        p += n
    out = bytearray(len(sels).to_bytes(8, "little")) + bytes([G - p])
    for k in range(0, len(sels), 2):
        out.append(sels[k] | ((sels[k + 1] << 4) if k + 1 < len(sels) else 0))
    while (base + len(out)) % (8 * E):
        out.append(0)
    q = 0
    for i in sels:
        for j in range(E):
            w = 0
            for k in range(nums[i]):
                w |= int(v[(q + k) * E + j]) << (k * bits[i])
            out += w.to_bytes(8, "little")
        q += nums[i]
    if G - q:
        R = G - q
        for j in range(E):
            w = 0
            for k in range(R):
                w |= int(v[(q + k) * E + j]) << (k * (64 // R))
            out += (w & (2**64 - 1)).to_bytes(8, "little")
    return bytes(out), sels


@pytest.mark.parametrize("E", [8, 4])
@pytest.mark.parametrize("kind", ["narrow", "mixed", "wide", "zeros", "short"])
def test_group_simple_against_model(oracle, E, kind):
    n = {"short": 5 * E, "zeros": 300 * E}.get(kind, 2000 * E)
    if kind == "narrow":
        v = rng_vals(1, n, 3)
    elif kind == "mixed":
        v = rng_vals(2, n, 30) >> (rng_vals(3, n, 5) % np.uint64(30))
    elif kind == "wide":
        v = rng_vals(4, n, 64)
    elif kind == "zeros":
        v = np.zeros(n, np.uint64)
    else:
        v = rng_vals(5, n, 10)
    f = group_simple_f(E)
    img = oracle.morph_batch_compress(f, v)
    want, sels = _gs_reference_model(v, E)
    assert img.tobytes() == want
    dec, used = oracle.morph_batch_decompress(f, img, n)
    assert used == len(img) and np.array_equal(dec, v)


@pytest.mark.parametrize("f", [static_vbp_padding_f(7, 8), k_wise_ns_f(2), group_simple_f(8)], ids=repr)
@pytest.mark.parametrize("n", [0, 1, 167, 4096 + 77])
def test_extra_formats_column_layout(oracle, f, n):
    v = rng_vals(n + 9, n, 7)
    col = oracle.morph(f, v)
    bs = f.block_size()
    nc = n // bs * bs
    main = oracle.morph_batch_compress(f, v[:nc]) if nc else np.zeros(0, np.uint8)
    if f.kind != 8 or nc:
        assert col.size_compr == len(main)
    assert np.array_equal(col.values(), v)
    if n - nc:
        assert col.count_uncompr() == n - nc


# ---------------------------------------------------------------- blockwise_individual_f
def _bwi_payload(oracle, code, blk):
    """One morph_batch call of the code's default format (rle_f: its pairs)."""
    from engine_amd.formats import rle_f
    if code == 71:
        return oracle.morph(rle_f(), blk).image()
    return oracle.morph_batch_compress(oracle.bwi_inner(code), blk)


@pytest.mark.parametrize("B", [1024, 2048, 3072])
@pytest.mark.parametrize("seed", [1, 2])
def test_bwi_layout(oracle, B, seed):
    """blockwise_individual.h:185-249: per block the code byte, zero padding
    to the next 64-byte boundary of the image (create_aligned_ptr with the
    64-byte vectors of the 8-lane style), then the default format's own
    morph_batch output; the rest raw behind the gap (column.h:182-196)."""
    from tests import bwi_data
    from engine_amd.formats import blockwise_individual_f
    nb, rest = 9, 123
    x = bwi_data.mixture(seed, nb, B, rest)
    allowed = [0, 7, 33, 64, 65, 66, 67, 69, 71]
    codes = bwi_data.fitting_codes(x, B, seed, allowed)
    f = blockwise_individual_f(B)
    c = oracle.morph_codes(f, x, codes)
    img = c.image()
    count, used, compr = c.meta()
    pos = 0
    for b in range(nb):
        assert img[pos] == codes[b]
        q = (pos + 1 + 63) // 64 * 64
        assert not img[pos + 1:q].any()
        pay = _bwi_payload(oracle, int(codes[b]), x[b * B:(b + 1) * B])
        assert img[q:q + len(pay)].tobytes() == pay.tobytes()
        pos = q + len(pay)
    assert pos == compr and count == nb * B + rest
    assert np.array_equal(oracle.block_codes(c), codes)
    assert np.array_equal(c.values(), x)
    r = oracle.rest_offset(f, compr)
    assert used == r + 8 * rest
    assert img[r:].view(np.uint64).tolist() == x[nb * B:].tolist()


def _bwi_sizes(oracle, blk):
    mbw = max(1, int(np.bitwise_or.reduce(blk)).bit_length())
    sizes = {c: len(_bwi_payload(oracle, c, blk)) for c in (65, 66, 67, 69)}
    return mbw, sizes


@pytest.mark.parametrize("kind", range(6))
def test_bwi_exhaustive_selectors(oracle, kind):
    """exhaustive_size_format_selector.h:80-141: best = the smallest of
    static_vbp at the maximum bit width (below 64 bits; uncompr_f
    otherwise) and dynamic_vbp, group_simple, delta and for in this order,
    a later one only when strictly smaller; worst starts from uncompr_f and
    takes a later one when not smaller.  Sizes here come from compressing
    the block in each format; the pick must be a size extremum."""
    from tests import bwi_data
    from engine_amd.formats import MS_BWI_SEL_BEST, MS_BWI_SEL_WORST, blockwise_individual_f
    B = 2048
    for s in range(3):
        blk = bwi_data.block_values(100 * kind + s, kind, B)
        mbw, sizes = _bwi_sizes(oracle, blk)
        order = [(mbw, mbw * B // 8) if mbw < 64 else (0, 8 * B)] + [(c, sizes[c]) for c in (65, 66, 67, 69)]
        best = oracle.bwi_choose(blockwise_individual_f(B, MS_BWI_SEL_BEST), blk)
        lo = min(sz for _, sz in order)
        assert best == next(c for c, sz in order if sz == lo)
        worst = oracle.bwi_choose(blockwise_individual_f(B, MS_BWI_SEL_WORST), blk)
        order_w = [(0, 8 * B)] + [(c, sizes[c]) for c in (65, 66, 67, 69)]
        hi = max(sz for _, sz in order_w)
        assert worst == [c for c, sz in order_w if sz == hi][-1]


def test_bwi_static_selector_and_refused_codes(oracle):
    from tests import bwi_data
    from engine_amd.formats import MS_BWI_SEL_STATIC, blockwise_individual_f
    x = bwi_data.mixture(5, 4, 1024, 10, kinds=2)
    for code in (67, 71, 64, 0):   # 71: rle pairs of random values take 16 bytes per value
        c = oracle.morph(blockwise_individual_f(1024, MS_BWI_SEL_STATIC, code), x)
        assert oracle.block_codes(c).tolist() == [code] * 4
        assert np.array_equal(c.values(), x)
    for bad in (68, 70, 72, 200):
        assert oracle.bwi_inner(bad) is None
        with pytest.raises(RuntimeError):
            oracle.morph_codes(blockwise_individual_f(1024), x, np.full(4, bad, np.uint8))


def test_bwi_prepare_for_random_access(oracle):
    """column.h:241-268 on blockwise_individual_f: the zero-padded rest is
    one more block at sizeCompr, aligned to the image."""
    from tests import bwi_data
    from engine_amd.formats import blockwise_individual_f
    B = 1024
    x = bwi_data.mixture(8, 3, B, 300)
    c = oracle.morph(blockwise_individual_f(B), x)
    compr = c.meta()[2]
    assert c.prepare_for_random_access()
    img = c.image()
    pad = np.concatenate([x[3 * B:], np.zeros(B - 300, np.uint64)])
    code = oracle.bwi_choose(blockwise_individual_f(B), pad)
    assert img[compr] == code
    q = (compr + 1 + 63) // 64 * 64
    pay = _bwi_payload(oracle, code, pad)
    assert img[q:q + len(pay)].tobytes() == pay.tobytes()
