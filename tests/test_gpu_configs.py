"""The benchmark configurations at their full sizes, through the C-ABI:

  * format cascade sweep: four uint64 columns of 2^27 values (dynamic_vbp
    random widths, delta+dynamic_vbp over sorted gaps, for+dynamic_vbp,
    rle); byte images equal the oracle's, round trips are exact and
    select+sum (single-column and batched) equals numpy;
  * two-column query plan over 2^28 rows: both select outputs and their
    intersect (two compressed position lists) equal the oracle's byte
    images, and the plan's sum (eager and replayed hipGraph) equals numpy;
  * 4096 uneven mixed-format columns: every batched sum (default and tuned
    shapes) equals numpy and the single-column path;
  * learned format selection over 1024 columns of 64 x 2^16 values: every
    column's sum equals the uncompressed sum on every path and the
    footprint is within 3% of the exact-best choice.
"""
import sys

import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen, hip
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_LT, default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f,
                                dynamic_vbp_f, rle_f, uncompr_f)
from tests import refmodel as rm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def same_image(g, o):
    img, meta = g.image()
    assert meta == o.meta(), (meta, o.meta())
    assert img.tobytes() == o.image().tobytes()


SWEEP = [("dynamic_vbp", lambda s, n: (datagen.dvbp_random_bw(s, n), 1 << 16), dynamic_vbp_f(512, 2048, 8)),
         ("delta", lambda s, n: (lambda x: (x, int(x[n // 2])))(datagen.sorted_geometric(s, n)),
          default_delta_dynamic_vbp_f(8)),
         ("for", lambda s, n: (lambda x: (x, int(x.min()) + (1 << 15)))(datagen.for_base_offset(s, n)),
          default_for_dynamic_vbp_f(8)),
         ("rle", lambda s, n: (datagen.rle_runs(s, n), 1 << 19), rle_f())]


@pytest.mark.parametrize("name,gen,f", SWEEP, ids=[s[0] for s in SWEEP])
def test_cascade_sweep_full_size(oracle, name, gen, f):
    n = 1 << 27
    x, val = gen(datagen.BASE_SEED + 1, n)
    say(f"sweep {name}: {n} values")
    col = ea.morph(f, ea.upload_values(x))
    same_image(col, oracle.morph(f, x))
    assert np.array_equal(col.values(), x)
    want = rm.masked_sum(x, MS_CMP_LT, val)
    assert int(ea.select_sum(MS_CMP_LT, val, col).get()[0]) == want
    b = ea.Batch()
    b.enqueue(col, MS_CMP_LT, val)
    b.enqueue(col, MS_CMP_EQ, int(x[n // 3]))
    res = ea.DeviceU64(2)
    b.run(res)
    assert res.get().tolist() == [want, rm.masked_sum(x, MS_CMP_EQ, int(x[n // 3]))]


def test_two_column_plan_full_size(oracle):
    n = 1 << 28
    seed = datagen.BASE_SEED + 2
    f = dynamic_vbp_f(512, 2048, 8)
    pos_f = default_delta_dynamic_vbp_f(8)
    fl = datagen.zipf_flags(seed, n)
    ky = datagen.uniform_bits(seed + 1, n, 20)
    say(f"plan: {n} rows generated")
    st = hip.Stream()
    cfl = ea.morph(f, ea.upload_values(fl, st.handle), st.handle)
    cky = ea.morph(f, ea.upload_values(ky, st.handle), st.handle)
    p1 = ea.select(MS_CMP_EQ, 3, pos_f, cfl, 8, st.handle)
    p2 = ea.select(MS_CMP_LT, 1 << 18, pos_f, cky, 8, st.handle)
    o1 = oracle.select(MS_CMP_EQ, 3, pos_f, oracle.from_values(fl), 8)
    same_image(p1, o1)
    o2 = oracle.select(MS_CMP_LT, 1 << 18, pos_f, oracle.from_values(ky), 8)
    same_image(p2, o2)
    say("plan: select outputs equal the oracle's")
    gi = ea.intersect_sorted(pos_f, p1, p2, st.handle)
    same_image(gi, oracle.intersect_sorted(pos_f, o1, o2))
    sel = (fl == 3) & (ky < (1 << 18))
    assert np.array_equal(gi.values(), np.flatnonzero(sel).astype(np.uint64))
    del o1, o2, gi, p1, p2
    say("plan: intersect of the two compressed lists equals the oracle's")
    ms = datagen.uniform_bits(seed + 2, n, 16)
    with np.errstate(over="ignore"):
        want = int(np.sum(ms[sel], dtype=np.uint64))
    cms = ea.morph(f, ea.upload_values(ms, st.handle), st.handle)
    del fl, ky, ms, sel

    def query(s):
        a = ea.select(MS_CMP_EQ, 3, pos_f, cfl, 8, s)
        b = ea.select(MS_CMP_LT, 1 << 18, pos_f, cky, 8, s)
        i = ea.intersect_sorted(pos_f, a, b, s)
        m = ea.project(uncompr_f(), cms, i, 8, s)
        return ea.agg_sum(m, s)
    assert int(query(st.handle).values(st.handle)[0]) == want
    plan = ea.Plan()
    with plan.record(st.handle):
        query(st.handle)
    with plan.record(st.handle):
        r = query(st.handle)
    for _ in range(3):
        plan.launch(st.handle)
        st.synchronize()
        assert int(r.values(st.handle)[0]) == want
    plan.free()


def test_uneven_columns_full_size():
    cols, expect = [], []
    for kind, f, x, val in datagen.mixed_columns(datagen.BASE_SEED + 3, 4096):
        cols.append((ea.morph(f, ea.upload_values(x)), val))
        expect.append(rm.masked_sum(x, MS_CMP_LT, val))
    say(f"uneven: {len(cols)} columns prepared")
    b = ea.Batch()
    for col, val in cols:
        b.enqueue(col, MS_CMP_LT, val)
    res = ea.DeviceU64(len(cols))
    b.run(res)
    assert res.get().tolist() == expect
    b.tune(res, reps=2)
    res.zero()
    b.run(res)
    assert res.get().tolist() == expect
    single = ea.DeviceU64(len(cols))
    for k, (col, val) in enumerate(cols):
        ea.select_sum(MS_CMP_LT, val, col, single.at(k))
    assert single.get().tolist() == expect


def test_learned_selection_full_size():
    from engine_amd import learn
    rep = learn.run_selection_workload(ncols=1024, chunks_per_col=64, steps=2, warmup=1, log=say)
    # every path's per-column sums are checked against the uncompressed
    # columns inside the workload (it raises on a difference)
    assert rep["chunks"] == 1024 * 64
    assert rep["learned_over_exact"] <= 1.03
