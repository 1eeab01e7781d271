"""Pins the CPU restatement of the learned format selection (oracle/chunks.py,
oracle/mlp.py): candidate sizes against the oracle's own byte images,
statistics against brute force, the generator against its definitions, the
MLP gradient against central differences."""
import numpy as np
import pytest

from oracle import chunks as ch
from oracle import mlp as om

SEED = 0x4D6F7270 + 5


def sample_chunks():
    t10, t64, t3 = ch.geo_table(10), ch.geo_table(64), ch.geo_table(3)
    return [
        ("uniform7", ch.gen_chunk(ch.GEN_UNIFORM, 7, 0, SEED, 0)),
        ("uniform64", ch.gen_chunk(ch.GEN_UNIFORM, 64, 0, SEED + 1, 0)),
        ("uniform0", ch.gen_chunk(ch.GEN_UNIFORM, 0, 0, SEED + 2, 0)),
        ("dvbp1-32", ch.gen_chunk(ch.GEN_DVBP, 1, 32, SEED + 3, 0)),
        ("dvbp40-64", ch.gen_chunk(ch.GEN_DVBP, 40, 64, SEED + 4, 0)),
        ("sorted10", ch.gen_chunk(ch.GEN_SORTED, 10, 0, SEED + 5, 123456789, t10)),
        ("sorted3_big", ch.gen_chunk(ch.GEN_SORTED, 3, 0, SEED + 6, 2**63 + 5, t3)),
        ("for16", ch.gen_chunk(ch.GEN_FOR, 16, 0, SEED + 7, (1 << 40) + 77)),
        ("for_wrap", ch.gen_chunk(ch.GEN_FOR, 8, 0, SEED + 8, 2**64 - 100)),
        ("rle64", ch.gen_chunk(ch.GEN_RLE, 64, 20, SEED + 9, 0, t64)),
        ("rle3", ch.gen_chunk(ch.GEN_RLE, 3, 2, SEED + 10, 0, t3)),
    ]


def candidate_formats(static_bw):
    from engine_amd.formats import (default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, dynamic_vbp_f, rle_f,
                                    static_vbp_f)
    return [static_vbp_f(static_bw, 8), dynamic_vbp_f(512, 64, 8), default_delta_dynamic_vbp_f(8),
            default_for_dynamic_vbp_f(8), rle_f()]


@pytest.mark.parametrize("name,x", sample_chunks(), ids=[n for n, _ in sample_chunks()])
def test_candidate_sizes_match_oracle_images(oracle, name, x):
    st, bw, dbw, fbw = ch.stats(x)
    want = ch.sizes(st)
    for k, f in enumerate(candidate_formats(int(st[ch.ST_MAX_BW]))):
        col = oracle.morph(f, x)
        assert col.meta()[1] == int(want[k]), (name, k, col.meta(), int(want[k]))
        assert np.array_equal(col.values(), x)


def test_stats_brute_force():
    x = ch.gen_chunk(ch.GEN_DVBP, 3, 20, SEED + 11, 0)
    x[5] = 0
    x[700:710] = 9
    st, bw, dbw, fbw = ch.stats(x)
    xs = [int(v) for v in x]
    M = (1 << 64) - 1
    for b in range(ch.BLOCKS):
        o = 0
        for v in xs[512 * b:512 * b + 512]:
            o |= v
        assert bw[b] == max(1, o.bit_length())
        o = 0
        for i in range(512 * b, 512 * b + 512):
            o |= (xs[i] - (xs[i - 8] if i >= 8 else 0)) & M
        assert dbw[b] == max(1, o.bit_length())
    runs = 1 + sum(1 for i in range(1, len(xs)) if xs[i] != xs[i - 1])
    assert int(st[ch.ST_RUNS]) == runs
    assert int(st[ch.ST_NONDEC]) == sum(1 for i in range(1, len(xs)) if xs[i] >= xs[i - 1])
    assert int(st[ch.ST_ZEROS]) == xs.count(0)
    assert int(st[ch.ST_SUM_BW2]) == sum(int(b) ** 2 for b in bw)


def test_for_refs_signed_lane_rule():
    x = np.full(1024, 5, np.uint64)
    x[3] = np.uint64(2**63 + 1)          # negative as int64: the lane minimum of lane 3
    r = ch.for_refs(x)
    assert int(r[0]) == 5                 # unsigned minimum over lanes {5, ..., 2^63+1}
    y = np.full(1024, 2**64 - 1, np.uint64)
    assert int(ch.for_refs(y)[0]) == 2**64 - 2


def test_generator_definitions():
    t10 = ch.geo_table(10)
    g = ch.geo(t10, ch.splitmix(SEED, np.arange(200000)))
    assert abs(g.mean() - 10.0) < 0.15 and g.min() >= 1
    assert ch.geo(ch.geo_table(1), np.array([0, 2**63], np.uint64)).tolist() == [1, 1]
    x = ch.gen_chunk(ch.GEN_DVBP, 5, 9, SEED, 0)
    st, bw, _, _ = ch.stats(x)
    blk = np.arange(ch.BLOCKS, dtype=np.uint64)
    drawn = 5 + ch.splitmix(SEED ^ 0xB1, blk) % np.uint64(5)
    assert np.array_equal(bw, drawn.astype(np.int64))
    s = ch.gen_chunk(ch.GEN_SORTED, 10, 0, SEED, 1000, t10)
    assert s[0] == 1000 and np.all(s[1:] > s[:-1])
    r = ch.gen_chunk(ch.GEN_RLE, 64, 20, SEED, 0, ch.geo_table(64))
    runs = 1 + np.count_nonzero(r[1:] != r[:-1])
    assert 700 < runs < 1400 and int(r.max()) < 2**20
    f = ch.gen_chunk(ch.GEN_FOR, 12, 0, SEED, 1 << 40)
    assert int(f.min()) >= 1 << 40 and int(f.max()) < (1 << 40) + 4096


def test_mlp_gradient_matches_central_differences():
    rng = np.random.default_rng(7)
    p = rng.normal(0, 0.3, om.NPARAM)
    X = rng.normal(0, 1, (13, om.IN))
    T = rng.normal(0, 1, (13, om.OUT))
    w = np.linspace(0.5, 1.5, om.OUT)
    L, g = om.loss_grad(p, X, T, w)
    assert abs(L - om.loss(p, X, T, w)) < 1e-12
    idx = list(range(0, om.NPARAM, 37)) + [om.NPARAM - 1, om.NPARAM - 11]
    num = om.numeric_grad(p, X, T, w, idx)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-7)
