"""Learned format selection on the GPU: the chunk generator, statistics,
features and exact sizes against the CPU restatement (bit-exact / float
tolerance), the MLP gradient against the float64 oracle, training, and the
selection's footprint and bit-exactness."""
import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen, hip, learn
from oracle import chunks as och
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu
SEED = 0x4D6F7270 + 6


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def edge_specs(tables):
    t2, t10, t1, t64 = tables(2), tables(10), tables(1), tables(64)
    return [
        (learn.GEN_UNIFORM, 0, 0, 0, SEED, 0), (learn.GEN_UNIFORM, 13, 0, 0, SEED + 1, 0),
        (learn.GEN_UNIFORM, 64, 0, 0, SEED + 2, 0),
        (learn.GEN_DVBP, 1, 32, 0, SEED + 3, 0), (learn.GEN_DVBP, 64, 64, 0, SEED + 4, 0),
        (learn.GEN_DVBP, 1, 1, 0, SEED + 5, 0),
        (learn.GEN_SORTED, 10, 0, t10, SEED + 6, 12345), (learn.GEN_SORTED, 2, 0, t2, SEED + 7, 2**64 - 50000),
        (learn.GEN_FOR, 16, 0, 0, SEED + 8, (1 << 40) + 3), (learn.GEN_FOR, 64, 0, 0, SEED + 9, 2**63),
        (learn.GEN_RLE, 64, 20, t64, SEED + 10, 0), (learn.GEN_RLE, 1, 64, t1, SEED + 11, 0),
        (learn.GEN_RLE, 2, 1, t2, SEED + 12, 0),
    ]


def oracle_chunk(spec, tables):
    kind, a, b, tab, seed, base = spec
    thr = tables.tabs[tab] if kind in (learn.GEN_SORTED, learn.GEN_RLE) else None
    return och.gen_chunk(kind, a, b, seed, base, thr)


def test_generator_stats_features_sizes_match_oracle():
    tables = learn.GeoTables()
    specs = edge_specs(tables)
    n = len(specs)
    d_vals = hip.DeviceArray(n * learn.CHUNK, np.uint64)
    learn.gen_chunks(specs, tables, d_vals)
    got = d_vals.numpy().reshape(n, learn.CHUNK)
    cost = np.abs(np.sin(np.arange(5 * 65, dtype=np.float64))).astype(np.float32).reshape(5, 65) + 0.1
    d_st = hip.DeviceArray(n * learn.NSTAT, np.uint64)
    d_tl = hip.DeviceArray(n * learn.NCAND, np.float32)
    learn.chunk_stats(d_vals, n, d_st, cost, d_tl)
    d_f = hip.DeviceArray(n * learn.NFEAT, np.float32)
    d_sz = hip.DeviceArray(n * learn.NCAND, np.uint64)
    learn.chunk_features(d_st, n, d_f, d_sz)
    st = d_st.numpy().reshape(n, -1)
    feats = d_f.numpy().reshape(n, -1)
    sz = d_sz.numpy().reshape(n, -1)
    tl = d_tl.numpy().reshape(n, -1)
    for i, sp in enumerate(specs):
        x = oracle_chunk(sp, tables)
        assert np.array_equal(got[i], x), (i, sp)
        ost, bw, dbw, fbw = och.stats(x)
        assert np.array_equal(st[i], ost), (i, sp, st[i], ost)
        assert np.array_equal(sz[i], och.sizes(ost)), i
        assert np.allclose(feats[i], och.features(ost), atol=2e-6), (i, feats[i], och.features(ost))
        assert np.allclose(tl[i], och.time_labels(ost, bw, dbw, fbw, cost), rtol=1e-5), i


def test_mlp_gradient_matches_float64_oracle():
    rng = np.random.default_rng(3)
    n = 1000
    X = rng.normal(0, 1, (n, 24)).astype(np.float32)
    Y = rng.normal(0, 1, (n, 10)).astype(np.float32)
    w = np.linspace(0.5, 2.0, 10).astype(np.float32)
    m = learn.MLP(seed=99)
    p = m.params
    L, g = m.loss_grad(hip.DeviceArray.from_numpy(X.reshape(-1)), hip.DeviceArray.from_numpy(Y.reshape(-1)), n, w)
    Lo, go = omlp.loss_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), w.astype(np.float64))
    assert abs(L - Lo) <= 1e-4 * abs(Lo)
    scale = np.abs(go).max()
    assert np.max(np.abs(g - go)) <= 2e-4 * scale, np.max(np.abs(g - go)) / scale
    # a parameter change moves the GPU loss as the oracle predicts
    p2 = p + 1e-3 * np.sign(go).astype(np.float32)
    m.params = p2
    L2, _ = m.loss_grad(hip.DeviceArray.from_numpy(X.reshape(-1)), hip.DeviceArray.from_numpy(Y.reshape(-1)), n, w)
    assert abs(L2 - omlp.loss(p2.astype(np.float64), X, Y, w)) <= 1e-4 * abs(L2)


def synthetic_cost():
    """A fixed decode-cost table (ns per unit by bit width) for tests that
    must not depend on timing."""
    bw = np.arange(65, dtype=np.float32)
    cost = np.zeros((5, 65), np.float32)
    cost[0] = 0.10 + 0.010 * bw
    cost[1] = 0.14 + 0.010 * bw
    cost[2] = 0.30 + 0.012 * bw
    cost[3] = 0.20 + 0.011 * bw
    cost[4, 0], cost[4, 1] = 5.0, 0.02
    return cost


def test_training_and_selection():
    tables = learn.GeoTables()
    cost = synthetic_cost()
    ntrain = 8192
    d_x, d_y, sizes, times = learn.label(learn.training_specs(SEED, ntrain, tables), tables, cost)
    m = learn.MLP(seed=5)
    losses = m.train(d_x, d_y, ntrain, np.ones(10, np.float32), epochs=40, batch=1024, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0], losses
    # held-out chunks of the selection workload
    specs = [s for c in range(8) for s in learn.mixture_specs(SEED + 1, c, 64, tables)]
    n = len(specs)
    d_f, _, esz, etm = learn.label(specs, tables, cost)
    d_sz = hip.DeviceArray.from_numpy(esz.reshape(-1))
    pred = m.predict(d_f, n)
    choice, fell_back = learn.select_formats(pred, d_sz, n)
    best = esz.min(axis=1).astype(np.float64)
    picked = esz[np.arange(n), choice].astype(np.float64)
    assert picked.sum() <= 1.03 * best.sum()
    assert np.all(picked <= 1.03 * best + 1e-9)
    assert fell_back.mean() < 0.25, fell_back.mean()
    # with this cost table the exact-smallest pick is (nearly) the fastest:
    # the time-aware pick must not be slower in total
    exact_choice = esz.argmin(axis=1)
    t_learned = etm[np.arange(n), choice].sum()
    t_exact = etm[np.arange(n), exact_choice].sum()
    assert t_learned <= t_exact * 1.02, (t_learned, t_exact)
    # rule 1: exact sizes admit the candidates within 3% of the smallest,
    # the predicted times choose among them (a 15% predicted gain to switch)
    choice1, kept = learn.select_formats(pred, d_sz, n, rule=1)
    assert np.all(esz[np.arange(n), choice1] <= 1.03 * best + 1e-9)
    assert np.array_equal(kept, choice1 == exact_choice)
    pt = np.exp2(pred.numpy(n * 10).reshape(n, 10)[:, 5:].astype(np.float64) * learn.TIME_SCALE + learn.TIME_OFF)
    sw = np.flatnonzero(~kept)
    assert np.all(pt[sw, choice1[sw]] * 1.15 <= pt[sw, exact_choice[sw]] * (1 + 1e-4))
    t_rule1 = etm[np.arange(n), choice1].sum()
    assert t_rule1 <= t_exact * 1.01, (t_rule1, t_exact)


def test_network_gain_over_exact_sizing():
    """What the trained network adds to exact sizing: it predicts decode
    time, so where a candidate within 5% of the smallest size decodes much
    faster (here: for_f is made 3x cheaper per unit than dynamic_vbp), the
    learned pick takes it; exact sizing alone cannot see the difference."""
    tables = learn.GeoTables()
    cost = synthetic_cost()
    cost[3] = cost[1] / 3.0
    ntrain = 16384
    d_x, d_y, _, _ = learn.label(learn.training_specs(SEED + 7, ntrain, tables), tables, cost)
    m = learn.MLP(seed=11)
    for e, lr in ((20, 3e-3), (20, 1e-3)):
        m.train(d_x, d_y, ntrain, np.ones(10, np.float32), epochs=e, batch=1024, lr=lr)
    specs = [s for c in range(8) for s in learn.mixture_specs(SEED + 8, c, 64, tables)]
    n = len(specs)
    d_f, _, esz, etm = learn.label(specs, tables, cost)
    # dynamic_vbp<512, 64, 8> and for_f over it differ by about 3.5% in size
    # on uniform chunks (for_f's reference pages and per-call meta), so the
    # tolerances here are 5%
    choice, _ = learn.select_formats(m.predict(d_f, n), hip.DeviceArray.from_numpy(esz.reshape(-1)), n,
                                     size_tol=0.05, exact_tol=0.05)
    exact_choice = esz.argmin(axis=1)
    rows = np.arange(n)
    t_learned, t_exact = etm[rows, choice].sum(), etm[rows, exact_choice].sum()
    s_learned, s_exact = esz[rows, choice].sum(), esz[rows, exact_choice].sum()
    assert s_learned <= 1.05 * s_exact
    assert t_learned < 0.95 * t_exact, (t_learned, t_exact)


def test_selected_chunks_are_bit_exact():
    tables = learn.GeoTables()
    specs = learn.mixture_specs(SEED + 2, 0, 40, tables)
    n = len(specs)
    d_vals = hip.DeviceArray(n * learn.CHUNK, np.uint64)
    learn.gen_chunks(specs, tables, d_vals)
    host = d_vals.numpy().reshape(n, learn.CHUNK)
    d_st = hip.DeviceArray(n * learn.NSTAT, np.uint64)
    learn.chunk_stats(d_vals, n, d_st)
    st = d_st.numpy().reshape(n, -1)
    b = ea.Batch()
    want = []
    for i in range(n):
        for cand in range(5):
            f = learn.candidate_format(cand, int(st[i, learn.ST_MAX_BW]))
            col = ea.morph(f, ea.from_device(d_vals.ptr(i * learn.CHUNK), learn.CHUNK))
            if i % 8 == 0:
                assert np.array_equal(col.values(), host[i])
            val = int(np.median(host[i]))
            b.enqueue(col, 2, val)
            with np.errstate(over="ignore"):
                want.append(int(np.sum(host[i][host[i] < np.uint64(val)], dtype=np.uint64)))
    res = ea.DeviceU64(len(want))
    b.run(res)
    assert res.get().tolist() == want


def test_selection_workload_small():
    """The whole pipeline at a small scale: every path's per-column sums
    equal the uncompressed sums (checked inside), footprint within 3% of the
    exact-best choice."""
    rep = learn.run_selection_workload(ncols=6, chunks_per_col=8, ntrain=8192, epochs=20, steps=2, warmup=1,
                                       cost=synthetic_cost())
    assert rep["learned_over_exact"] <= 1.03
    assert rep["footprint_bytes"]["learned"] <= rep["footprint_bytes"]["plain"] * 1.03
    assert rep["chunks"] == 48


def test_view_times_follow_bytes():
    """The timed run: one time per view (a block of a blockwise_individual_f
    entry, or a whole entry), wider blocks take longer, and the run's sums
    are the ordinary ones."""
    from engine_amd.formats import blockwise_individual_f
    B = 1 << 16
    narrow = datagen.uniform_bits(SEED + 21, 16 * B, 2)
    wide = datagen.uniform_bits(SEED + 22, 16 * B, 48)
    x = np.concatenate([narrow, wide])
    codes = np.array([2] * 16 + [48] * 16, np.uint8)
    col = ea.morph_blockwise(blockwise_individual_f(B), ea.upload_values(x), codes)
    plain = ea.upload_values(narrow[:5000])
    b = ea.Batch()
    b.enqueue(col, 2, 1 << 40)
    b.enqueue(plain, 2, 3)
    res = ea.DeviceU64(2)
    t = b.view_times(res, reps=5)
    assert len(t) == 33 and (t > 0).all()
    assert t[16:32].mean() > t[:16].mean()
    got = res.get()
    assert int(got[0]) == int(np.sum(x[x < np.uint64(1 << 40)], dtype=np.uint64))
    assert int(got[1]) == int(np.sum(narrow[:5000][narrow[:5000] < 3], dtype=np.uint64))


def test_measured_labels():
    """Labels from compressing and timing: the same exact sizes as the
    closed forms, positive times, and wider chunks slower in the formats
    whose size follows the values' width (static_vbp, dynamic_vbp, for_f;
    the deltas of uniform values wrap to 64 bits at any width)."""
    tables = learn.GeoTables()
    specs = [(learn.GEN_UNIFORM, bw, 0, 0, SEED + bw + 7 * k, 0) for k in range(4) for bw in (2, 2, 40, 40)]
    d_x, d_y, sizes, times = learn.label_measured(specs, tables, batch=8, reps=5)
    _, _, sizes_t, _ = learn.label(specs, tables, synthetic_cost())
    assert np.array_equal(sizes, sizes_t)
    assert times.shape == (16, learn.NCAND) and (times > 0).all()
    narrow, wide = times[0::4] + times[1::4], times[2::4] + times[3::4]
    for k in (0, 1, 3):
        assert wide[:, k].mean() > narrow[:, k].mean(), k
