"""Parity of the HIP path (through the C-ABI) with the CPU oracle:
identical byte images and meta data for identical format parameters, and
bit-exact values, positions and sums."""
import json
import os

import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_GE, MS_CMP_GT, MS_CMP_LE, MS_CMP_LT, MS_CMP_NE, MS_ERR_SIZE, MS_OP_ADD,
                                default_delta_dynamic_vbp_f, default_dynamic_vbp_f, default_for_dynamic_vbp_f,
                                delta_f, dynamic_vbp_f, for_f, rle_f, static_vbp_f, uncompr_f)
from tests import refmodel as rm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "reference_kat.json")))

FORMATS = [
    static_vbp_f(12, 1), static_vbp_f(7, 8), static_vbp_f(33, 8), static_vbp_f(64, 8), static_vbp_f(5, 4),
    dynamic_vbp_f(512, 2048, 8), default_dynamic_vbp_f(8), dynamic_vbp_f(1024, 64, 8),
    default_dynamic_vbp_f(4), default_dynamic_vbp_f(1),
    default_delta_dynamic_vbp_f(8), delta_f(1024, 8, static_vbp_f(40, 8)), delta_f(512, 4, default_dynamic_vbp_f(4)),
    default_for_dynamic_vbp_f(8), for_f(1024, 8, uncompr_f()), for_f(2048, 16, dynamic_vbp_f(1024, 64, 8)),
    rle_f(),
]
FMT_IDS = [repr(f) for f in FORMATS]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def data_for(f, n, seed):
    if f.kind == 1:
        return datagen.uniform_bits(seed, n, f.bw)
    if f.kind == 3:
        return datagen.sorted_geometric(seed, n)
    if f.kind == 4:
        v = datagen.for_base_offset(seed, n)
        if n > 100:
            v[77] = np.uint64(2**63 + 11)     # signed lane minimum of the 8-lane reference value
        return v
    if f.kind == 5:
        return datagen.rle_runs(seed, n, mean_len=9)
    if n >= 512:
        return datagen.dvbp_random_bw(seed, n // 512 * 512, lo=1, hi=40).tolist() + \
            datagen.uniform_bits(seed + 1, n % 512, 20).tolist()
    return datagen.uniform_bits(seed, n, 20)


def assert_same_column(gcol, ocol):
    img, meta = gcol.image()
    assert meta == ocol.meta(), (meta, ocol.meta())
    oimg = ocol.image()
    assert img.tobytes() == oimg.tobytes()


# ---------------------------------------------------------------- morph
@pytest.mark.parametrize("f", FORMATS, ids=FMT_IDS)
@pytest.mark.parametrize("n", [0, 1, 100, 512, 4096 + 333, 70000])
def test_morph_compress_byte_identical(oracle, f, n):
    v = np.asarray(data_for(f, n, n + 7), dtype=np.uint64)
    g = ea.morph(f, ea.upload_values(v))
    o = oracle.morph(f, v)
    assert_same_column(g, o)
    assert np.array_equal(g.values(), v)


@pytest.mark.parametrize("f", FORMATS, ids=FMT_IDS)
def test_upload_oracle_image_then_decompress(oracle, f):
    n = 3 * 8192 + 77
    v = np.asarray(data_for(f, n, 3), dtype=np.uint64)
    o = oracle.morph(f, v)
    g = ea.upload_image(f, o.image(), *o.meta())
    assert np.array_equal(g.values(), v)
    assert np.array_equal(ea.morph(f, g).values(), v)      # same-format morph copies


# ------------------------------------------------------------ select_sum
CMPS = [MS_CMP_EQ, MS_CMP_NE, MS_CMP_LT, MS_CMP_LE, MS_CMP_GT, MS_CMP_GE]


@pytest.mark.parametrize("f", FORMATS + [uncompr_f()], ids=FMT_IDS + ["uncompr_f"])
def test_select_sum_all_formats(oracle, f):
    n = 5 * 4096 + 45
    v = np.asarray(data_for(f, n, 11), dtype=np.uint64)
    col = ea.morph(f, ea.upload_values(v)) if f.kind else ea.upload_values(v)
    ocol = oracle.morph(f, v)
    for val in [int(v[n // 2]), int(v[3]) + 1, 0, 2**64 - 1]:
        for cmp in CMPS:
            got = int(ea.select_sum(cmp, val, col).get()[0])
            assert got == oracle.select_sum(cmp, val, ocol) == rm.masked_sum(v, cmp, val), (cmp, val)
    s = ea.agg_sum(col)
    o = oracle.agg_sum(ocol)
    assert_same_column(s, o)


# ---------------------------------------------------------------- select
OUT_FORMATS = [uncompr_f(), default_delta_dynamic_vbp_f(8), default_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8),
               default_for_dynamic_vbp_f(8), static_vbp_f(20, 8)]
OUT_IDS = [repr(f) for f in OUT_FORMATS]


@pytest.mark.parametrize("lanes", [1, 8])
@pytest.mark.parametrize("out_f", OUT_FORMATS, ids=OUT_IDS)
@pytest.mark.parametrize("in_f", [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8),
                                  default_delta_dynamic_vbp_f(8), default_for_dynamic_vbp_f(8), rle_f()],
                         ids=lambda f: repr(f))
def test_select_byte_identical(oracle, lanes, out_f, in_f):
    n = 2 * 8192 + 517                     # rest of 517 values: 512 vectorised + 5 scalar with 8 lanes
    v = np.asarray(data_for(in_f, n, 5) if in_f.kind else datagen.uniform_bits(5, n, 12), dtype=np.uint64)
    col = ea.morph(in_f, ea.upload_values(v)) if in_f.kind else ea.upload_values(v)
    ocol = oracle.morph(in_f, v)
    for cmp, val in [(MS_CMP_LT, int(np.median(v))), (MS_CMP_EQ, int(v[-2])), (MS_CMP_GT, 2**64 - 1)]:
        g = ea.select(cmp, val, out_f, col, lanes)
        o = oracle.select(cmp, val, out_f, ocol, lanes)
        assert_same_column(g, o)
        assert np.array_equal(g.values(), rm.select_positions(v, cmp, val))


# select from dynamic_vbp / for_f inputs into the block formats packs its
# output straight from the count pass's ballots: densities from a handful
# of rows to all of them (units walked past empty, lanes full), delta steps
# equal to and below the lane count, and inner blocks of 512 to 2048
@pytest.mark.parametrize("out_f", [default_delta_dynamic_vbp_f(8), delta_f(1024, 4, dynamic_vbp_f(512, 64, 4)),
                                   dynamic_vbp_f(2048, 64, 8), delta_f(2048, 8, dynamic_vbp_f(2048, 64, 8)),
                                   for_f(1024, 8, dynamic_vbp_f(512, 64, 8))], ids=repr)
@pytest.mark.parametrize("in_f", [dynamic_vbp_f(512, 2048, 8), default_for_dynamic_vbp_f(8)], ids=repr)
def test_select_positions_from_ballots(oracle, out_f, in_f):
    n = 3 * 65536 + 777
    v = datagen.uniform_bits(21, n, 20)
    col = ea.morph(in_f, ea.upload_values(v))
    ocol = oracle.morph(in_f, v)
    for thr in (0, 40, 3000, 300000, 950000, 1 << 20):
        for lanes in (1, 8):
            g = ea.select(MS_CMP_LT, thr, out_f, col, lanes)
            o = oracle.select(MS_CMP_LT, thr, out_f, ocol, lanes)
            assert_same_column(g, o)
            assert np.array_equal(g.values(), np.flatnonzero(v < thr).astype(np.uint64)), (thr, lanes)


@pytest.mark.parametrize("in_f", [dynamic_vbp_f(512, 2048, 8), default_for_dynamic_vbp_f(8)], ids=repr)
def test_select_positions_from_ballots_small_and_ragged(oracle, in_f):
    """The ballot-fed emit at the edges: empty and one-value inputs, counts
    around a unit, a delta tile and a write-iterator flush, with and
    without a scalar remainder."""
    for n in (0, 1, 5, 511, 512, 513, 1023, 2049, 4095, 4097):
        v = datagen.uniform_bits(30 + n, n, 9)
        col = ea.morph(in_f, ea.upload_values(v))
        ocol = oracle.morph(in_f, v)
        for out_f in (default_delta_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8)):
            for thr, lanes in ((0, 8), (300, 1), (300, 8), (1 << 9, 8)):
                g = ea.select(MS_CMP_LT, thr, out_f, col, lanes)
                o = oracle.select(MS_CMP_LT, thr, out_f, ocol, lanes)
                assert_same_column(g, o)
                assert np.array_equal(g.values(), np.flatnonzero(v < thr).astype(np.uint64)), (n, thr, lanes)


@pytest.mark.parametrize("n", [1, 5, 13, 100])
@pytest.mark.parametrize("out_f", OUT_FORMATS, ids=OUT_IDS)
def test_select_tiny_inputs(oracle, n, out_f):
    v = datagen.uniform_bits(n, n, 3)
    for lanes in (1, 8):
        for val in (0, 3, 8):
            g = ea.select(MS_CMP_LT, val, out_f, ea.upload_values(v), lanes)
            o = oracle.select(MS_CMP_LT, val, out_f, oracle.from_values(v), lanes)
            assert_same_column(g, o)


# --------------------------------------------------------------- project
@pytest.mark.parametrize("data_f", [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8),
                                    default_for_dynamic_vbp_f(8), default_delta_dynamic_vbp_f(8), rle_f()],
                         ids=lambda f: repr(f))
@pytest.mark.parametrize("pos_f", [uncompr_f(), default_delta_dynamic_vbp_f(8)], ids=lambda f: repr(f))
@pytest.mark.parametrize("out_f", [uncompr_f(), default_dynamic_vbp_f(8)], ids=lambda f: repr(f))
def test_project_byte_identical(oracle, data_f, pos_f, out_f):
    n = 8192 * 3 + 41
    v = np.asarray(data_for(data_f, n, 9), dtype=np.uint64)
    p = np.flatnonzero(datagen.uniform_bits(13, n, 4) < 6).astype(np.uint64)
    dcol = ea.morph(data_f, ea.upload_values(v)) if data_f.kind else ea.upload_values(v)
    pcol = ea.morph(pos_f, ea.upload_values(p)) if pos_f.kind else ea.upload_values(p)
    g = ea.project(out_f, dcol, pcol, 8)
    o = oracle.project(out_f, oracle.morph(data_f, v), oracle.morph(pos_f, p), 8)
    assert_same_column(g, o)
    assert np.array_equal(g.values(), v[p])
    g.check()


@pytest.mark.parametrize("out_f", [for_f(512, 1024, dynamic_vbp_f(512, 64, 8)), for_f(512, 8, dynamic_vbp_f(512, 64, 8))],
                         ids=repr)
def test_project_output_fits_per_flush_pages(oracle, out_f):
    """A for_f output whose reference page is larger than a 2048-value
    flush: every flush starts a fresh page, so the output needs far more
    than a column-level size bound.  It must fit its allocation, report no
    capacity error and equal the oracle byte for byte."""
    n = 1 << 19
    v = datagen.uniform_bits(19, n, 64)
    v[::512] |= np.uint64(1 << 63)
    p = np.arange(n, dtype=np.uint64)
    g = ea.project(out_f, ea.upload_values(v), ea.upload_values(p), 8)
    g.check()
    m = g.meta()
    assert m.size_used <= m.size_alloc
    assert_same_column(g, oracle.project(out_f, oracle.from_values(v), oracle.from_values(p), 8))
    assert np.array_equal(g.values(), v)


def test_project_out_of_range_reports():
    d = ea.upload_values(np.arange(10, dtype=np.uint64))
    p = ea.upload_values(np.array([1, 2, 10], np.uint64))
    g = ea.project(uncompr_f(), d, p, 8)
    with pytest.raises(ea.MsError) as e:
        g.check()
    assert e.value.code == 7


# ------------------------------------------------------- set operators
@pytest.mark.parametrize("out_f", [uncompr_f(), default_delta_dynamic_vbp_f(8)], ids=lambda f: repr(f))
def test_intersect_merge_calc(oracle, out_f):
    a = np.sort(datagen.uniform_bits(1, 30000, 15))
    b = np.sort(datagen.uniform_bits(2, 25000, 15))
    fa = default_delta_dynamic_vbp_f(8)
    ga, gb = ea.morph(fa, ea.upload_values(a)), ea.upload_values(b)
    oa, ob = oracle.morph(fa, a), oracle.from_values(b)
    assert_same_column(ea.intersect_sorted(out_f, ga, gb), oracle.intersect_sorted(out_f, oa, ob))
    assert_same_column(ea.merge_sorted(out_f, ga, gb), oracle.merge_sorted(out_f, oa, ob))
    assert np.array_equal(ea.intersect_sorted(out_f, ga, gb).values(), rm.multiset_intersect(a, b))
    assert np.array_equal(ea.merge_sorted(out_f, gb, ga).values(), rm.multiset_union(a, b))
    c = datagen.uniform_bits(3, 30000, 64)
    gc = ea.upload_values(c)
    for op in range(10):
        g = ea.calc_binary(op, out_f, ga, gc)
        o = oracle.calc_binary(op, out_f, oa, oracle.from_values(c))
        assert_same_column(g, o)
    with pytest.raises(ea.MsError) as e:
        ea.calc_binary(MS_OP_ADD, out_f, ga, gb)
    assert e.value.code == 4


@pytest.mark.parametrize("fb", [default_delta_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8)], ids=repr)
def test_intersect_merge_both_compressed(oracle, fb):
    """Both operands compressed position lists (the shape of a two-column
    plan): select outputs in delta+dynamic_vbp intersected with a second
    compressed list, written compressed; byte images equal the oracle's."""
    n = 300000
    x = datagen.uniform_bits(31, n, 20)
    y = datagen.uniform_bits(32, n, 3)
    fa = default_delta_dynamic_vbp_f(8)
    gx, gy = ea.upload_values(x), ea.upload_values(y)
    ox, oy = oracle.from_values(x), oracle.from_values(y)
    pa = ea.select(MS_CMP_LT, 1 << 18, fa, gx, 8)
    pb = ea.select(MS_CMP_EQ, 3, fb, gy, 8)
    oa = oracle.select(MS_CMP_LT, 1 << 18, fa, ox, 8)
    ob = oracle.select(MS_CMP_EQ, 3, fb, oy, 8)
    assert_same_column(pa, oa)
    assert_same_column(pb, ob)
    a = np.flatnonzero(x < (1 << 18)).astype(np.uint64)
    b = np.flatnonzero(y == 3).astype(np.uint64)
    for out_f in (fa, uncompr_f()):
        gi = ea.intersect_sorted(out_f, pa, pb)
        assert_same_column(gi, oracle.intersect_sorted(out_f, oa, ob))
        assert np.array_equal(gi.values(), rm.multiset_intersect(a, b))
        gm = ea.merge_sorted(out_f, pb, pa)
        assert_same_column(gm, oracle.merge_sorted(out_f, ob, oa))
        assert np.array_equal(gm.values(), rm.multiset_union(b, a))


@pytest.mark.parametrize("na,nb,bits", [(3000, 200000, 20), (5000, 5000, 6), (70000, 3, 17), (1, 1024, 4)])
def test_intersect_shapes(na, nb, bits):
    """Chunks of the left list whose range of the right list is larger than
    one workgroup stages (dense right side), runs of equal values across
    chunk boundaries, a tiny right side, a single left value."""
    a = np.sort(datagen.uniform_bits(10 + na, na, bits))
    b = np.sort(datagen.uniform_bits(20 + nb, nb, bits))
    got = ea.intersect_sorted(uncompr_f(), ea.upload_values(a), ea.upload_values(b)).values()
    assert np.array_equal(got, rm.multiset_intersect(a, b))
    got = ea.intersect_sorted(uncompr_f(), ea.upload_values(b), ea.upload_values(a)).values()
    assert np.array_equal(got, rm.multiset_intersect(b, a))


# ------------------------------------------------------------------ KATs
def test_reference_kats_on_gpu():
    k = KAT["select_less"]
    out = ea.select(MS_CMP_LT, k["val"], uncompr_f(), ea.upload_values(k["in"]), 1)
    assert out.values().tolist() == k["out"]
    k = KAT["project"]
    out = ea.project(uncompr_f(), ea.upload_values(k["data"]), ea.upload_values(k["pos"]), 1)
    assert out.values().tolist() == k["out"]
    k = KAT["agg_sum"]
    assert ea.agg_sum(ea.upload_values(k["in"])).values().tolist() == k["out"]
    k = KAT["agg_sum_1_to_100"]
    assert ea.agg_sum(ea.upload_values(np.arange(1, 101))).values().tolist() == k["out"]
    k = KAT["intersect_sorted"]
    out = ea.intersect_sorted(uncompr_f(), ea.upload_values(k["l"]), ea.upload_values(k["r"]))
    assert out.values().tolist() == k["out"]
    k = KAT["merge_sorted"]
    out = ea.merge_sorted(uncompr_f(), ea.upload_values(k["l"]), ea.upload_values(k["r"]))
    assert out.values().tolist() == k["out"]
    k = KAT["calc_binary_plus"]
    out = ea.calc_binary(MS_OP_ADD, uncompr_f(), ea.upload_values(k["l"]), ea.upload_values(k["r"]))
    assert out.values().tolist() == k["out"]


# ------------------------------------------------ full-size properties
def test_config_sizes_round_trip_and_sums():
    """2^24-value columns of every config-2 layout: decode(encode(x)) == x
    and the fused select + sum equals numpy's, at a size the oracle would
    take long for."""
    n = 1 << 24
    cases = [
        (static_vbp_f(12, 1), datagen.uniform_bits(0x4D6F7270, n, 12), 1 << 11),
        (dynamic_vbp_f(512, 2048, 8), datagen.dvbp_random_bw(0x4D6F7271, n), 1 << 16),
        (default_delta_dynamic_vbp_f(8), datagen.sorted_geometric(0x4D6F7271, n), 5 * n),
        (default_for_dynamic_vbp_f(8), datagen.for_base_offset(0x4D6F7271, n), 0),
        (rle_f(), datagen.rle_runs(0x4D6F7271, n), 1 << 19),
    ]
    for f, v, val in cases:
        if f.kind == 4:
            val = int(v.min()) + (1 << 15)
        col = ea.morph(f, ea.upload_values(v))
        assert np.array_equal(col.values(), v), repr(f)
        assert int(ea.select_sum(MS_CMP_LT, val, col).get()[0]) == rm.masked_sum(v, MS_CMP_LT, val), repr(f)


# ---------------------------------------------------- plans (hipGraphs)
def run_two_column_query(flag, key, measure, stream):
    """select(flag==3) -> delta+dvbp positions; intersect with
    select(key<2^18); project(measure); agg_sum."""
    pos_f = default_delta_dynamic_vbp_f(8)
    p1 = ea.select(MS_CMP_EQ, 3, pos_f, flag, 8, stream)
    p2 = ea.select(MS_CMP_LT, 1 << 18, pos_f, key, 8, stream)
    i = ea.intersect_sorted(pos_f, p1, p2, stream)
    m = ea.project(uncompr_f(), measure, i, 8, stream)
    return ea.agg_sum(m, stream), (p1, p2, i, m)


def test_plan_graph_replays_match_eager(oracle):
    from engine_amd import hip
    n = 3 * 65536 + 1234
    fl = datagen.zipf_flags(1, n)
    ky = datagen.uniform_bits(2, n, 20)
    ms = datagen.uniform_bits(3, n, 16)
    f = dynamic_vbp_f(512, 2048, 8)
    st = hip.Stream()
    flag, key, measure = (ea.morph(f, ea.upload_values(x), st.handle) for x in (fl, ky, ms))
    eager, _ = run_two_column_query(flag, key, measure, st.handle)
    want = eager.values(st.handle).tolist()
    sel = (fl == 3) & (ky < (1 << 18))
    with np.errstate(over="ignore"):
        assert want == [int(np.sum(ms[sel], dtype=np.uint64))]
    plan = ea.Plan()
    with plan.record(st.handle):                 # measuring pass
        r0, keep0 = run_two_column_query(flag, key, measure, st.handle)
    assert plan.arena_bytes > 0
    with plan.record(st.handle):                 # capture
        res, keep = run_two_column_query(flag, key, measure, st.handle)
    for rep in range(4):
        plan.launch(st.handle)
        st.synchronize()
        assert res.values(st.handle).tolist() == want, rep
        # eager work between replays
        other = int(ea.select_sum(MS_CMP_LT, 100, measure).get()[0])
        assert other == rm.masked_sum(ms, MS_CMP_LT, 100)
    del keep, res
    plan.free()


def test_calc_binary_in_a_plan():
    """calc_binary recorded into a plan: the count check runs on the device, so
    replays compute the same column as the eager call, and a count mismatch
    inside a plan shows as MS_ERR_SIZE in the output's status."""
    from engine_amd import hip
    n = 40_000 + 3
    x = datagen.uniform_bits(11, n, 30)
    y = datagen.uniform_bits(12, n, 30)
    st = hip.Stream()
    lx = ea.morph(default_dynamic_vbp_f(8), ea.upload_values(x), st.handle)
    ly = ea.upload_values(y)
    want = (x + y).tolist()
    assert ea.calc_binary(MS_OP_ADD, uncompr_f(), lx, ly, st.handle).values(st.handle).tolist() == want
    plan = ea.Plan()
    with plan.record(st.handle):
        ea.calc_binary(MS_OP_ADD, uncompr_f(), lx, ly, st.handle)
    with plan.record(st.handle):
        out = ea.calc_binary(MS_OP_ADD, uncompr_f(), lx, ly, st.handle)
    for rep in range(3):
        plan.launch(st.handle)
        st.synchronize()
        assert out.values(st.handle).tolist() == want, rep
    plan.free()
    short = ea.upload_values(y[:-1])
    with pytest.raises(ea.MsError):
        ea.calc_binary(MS_OP_ADD, uncompr_f(), lx, short, st.handle)
    p2 = ea.Plan(64 << 20)   # an arena given: the first recording captures (no eager measuring pass)
    with p2.record(st.handle):
        bad = ea.calc_binary(MS_OP_ADD, uncompr_f(), lx, short, st.handle)
    p2.launch(st.handle)
    st.synchronize()
    with pytest.raises(ea.MsError) as e:
        bad.check(st.handle)
    assert e.value.code == MS_ERR_SIZE
    p2.free()


PRA_FORMATS = [static_vbp_f(12, 1), static_vbp_f(7, 8), static_vbp_f(40, 8), dynamic_vbp_f(512, 2048, 8),
               default_dynamic_vbp_f(8), default_delta_dynamic_vbp_f(8), default_for_dynamic_vbp_f(8),
               delta_f(1024, 8, static_vbp_f(40, 8)), for_f(1024, 8, uncompr_f())]


@pytest.mark.parametrize("f", PRA_FORMATS, ids=repr)
@pytest.mark.parametrize("n", [0, 5, 4096, 4096 + 77, 100_000 + 333])
def test_prepare_for_random_access(oracle, f, n):
    """column<F>::prepare_for_random_access on the device: the image
    (including the compressed rest it puts into the gap) equals the
    oracle's after the same call; a second call, uncompr_f and a column
    without a rest do nothing; values, sums and gathers are unchanged."""
    x = np.ascontiguousarray(data_for(f, n, 90 + n % 11), dtype=np.uint64)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    want = o.prepare_for_random_access()
    assert g.prepare_for_random_access() == want
    assert_same_column(g, o)
    assert not g.prepare_for_random_access()
    assert np.array_equal(g.values(), x)
    if n:
        val = int(np.median(x))
        assert int(ea.select_sum(MS_CMP_LT, val, g).get()[0]) == rm.masked_sum(x, MS_CMP_LT, val)
        pos = np.array([0, n - 1, n // 2, max(0, n - 3)], dtype=np.uint64)
        got = ea.project(uncompr_f(), g, ea.upload_values(pos)).values()
        assert np.array_equal(got, x[pos.astype(np.int64)])
    # a same-format morph keeps the prepared image
    assert_same_column(ea.morph(f, g), o)


def test_prepare_for_random_access_uncompr_and_rle():
    u = ea.upload_values(np.arange(10, dtype=np.uint64))
    assert not u.prepare_for_random_access()
    r = ea.morph(rle_f(), u)
    with pytest.raises(ea.MsError):
        r.prepare_for_random_access()


@pytest.mark.parametrize("in_f", [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8),
                                  default_for_dynamic_vbp_f(8), default_delta_dynamic_vbp_f(8)], ids=repr)
@pytest.mark.parametrize("out_f", [uncompr_f(), default_delta_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8)],
                         ids=repr)
@pytest.mark.parametrize("lanes", [1, 8])
def test_between_parity(oracle, in_f, out_f, lanes):
    """between (two comparisons) on the device: byte images equal the
    oracle's for interval, empty and one-sided ranges."""
    n = 30_000 + 77
    x = np.ascontiguousarray(data_for(in_f, n, 33), dtype=np.uint64)
    g = ea.morph(in_f, ea.upload_values(x)) if in_f.kind else ea.upload_values(x)
    o = oracle.morph(in_f, x) if in_f.kind else oracle.from_values(x)
    q = np.quantile(x.astype(np.float64), [0.2, 0.7]).astype(np.uint64)
    for lo_cmp, lo, hi_cmp, hi in [(MS_CMP_GE, int(q[0]), MS_CMP_LE, int(q[1])),
                                   (MS_CMP_GT, int(q[0]), MS_CMP_LT, int(q[1])),
                                   (MS_CMP_GE, int(q[1]), MS_CMP_LE, int(q[0])),
                                   (MS_CMP_EQ, int(x[n // 2]), MS_CMP_GE, 0)]:
        assert_same_column(ea.between(lo_cmp, lo, hi_cmp, hi, out_f, g, lanes),
                           oracle.between(lo_cmp, lo, hi_cmp, hi, out_f, o, lanes))
    with pytest.raises(ea.MsError):
        ea.between(MS_CMP_NE, 1, MS_CMP_LE, 5, out_f, g, lanes)
