"""Batched select + sum (ms_batch_*): one persistent launch over many
columns of mixed formats and uneven lengths gives, per entry, exactly the
sum that numpy and the single-column path (ms_select_sum) give."""
import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_GE, MS_CMP_GT, MS_CMP_LE, MS_CMP_LT, MS_CMP_NE, BatchConfig,
                                default_delta_dynamic_vbp_f, default_dynamic_vbp_f, default_for_dynamic_vbp_f,
                                delta_f, dynamic_vbp_f, for_f, group_simple_f, rle_f, static_vbp_f, uncompr_f)
from tests import refmodel as rm

pytestmark = pytest.mark.gpu

CMPS = [MS_CMP_EQ, MS_CMP_NE, MS_CMP_LT, MS_CMP_LE, MS_CMP_GT, MS_CMP_GE]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def edge_columns():
    """Every batch item kind, rests, empty and tiny columns, wide codes and
    for_f blocks whose reference plus code range wraps."""
    seed = datagen.BASE_SEED + 40
    out = []
    for n in (0, 1, 511, 512, 513, 2047, 2048, 2049, 5000):
        out.append((static_vbp_f(9, 8), datagen.uniform_bits(seed + n, n, 9)))
        out.append((dynamic_vbp_f(512, 2048, 8), datagen.uniform_bits(seed + n + 1, n, 20)))
        out.append((default_delta_dynamic_vbp_f(8), datagen.sorted_geometric(seed + n + 2, n)))
        out.append((default_for_dynamic_vbp_f(8), datagen.for_base_offset(seed + n + 3, n)))
        out.append((rle_f(), datagen.rle_runs(seed + n + 4, n, mean_len=7)))
        out.append((uncompr_f(), datagen.uniform_bits(seed + n + 5, n, 30)))
    for bw in (1, 31, 32, 33, 47, 64):
        out.append((static_vbp_f(bw, 8), datagen.uniform_bits(seed + bw, 4096 + 77, bw)))
    out.append((dynamic_vbp_f(512, 2048, 8), datagen.dvbp_random_bw(seed, 8192, lo=1, hi=64)))
    x = datagen.uniform_bits(seed + 9, 3000, 8) + np.uint64(2**64 - 200)        # wrapping for_f blocks
    out.append((default_for_dynamic_vbp_f(8), x))
    x = datagen.for_base_offset(seed + 10, 9000)
    x[100] = np.uint64(2**63 + 5)
    out.append((default_for_dynamic_vbp_f(8), x))
    # layouts without a unit decoder (read value by value)
    out.append((static_vbp_f(5, 4), datagen.uniform_bits(seed + 11, 3000, 5)))
    out.append((default_dynamic_vbp_f(1), datagen.uniform_bits(seed + 12, 3000, 17)))
    out.append((for_f(1024, 8, uncompr_f()), datagen.for_base_offset(seed + 13, 5000)))
    out.append((dynamic_vbp_f(1024, 64, 8), datagen.uniform_bits(seed + 14, 7000, 13)))
    # delta_f layouts read difference by difference: steps 1, 2, 4, 8 over
    # other inner layouts, partial tiles and rests
    for k, n in enumerate((1, 2047, 2048, 2049, 9000)):
        for f in (default_delta_dynamic_vbp_f(4), default_delta_dynamic_vbp_f(2), default_delta_dynamic_vbp_f(1),
                  delta_f(1024, 8, static_vbp_f(24, 8)), delta_f(1024, 4, uncompr_f())):
            out.append((f, datagen.sorted_geometric(seed + 20 + k, n)))
    return out


def make_batch(cols_vals, cmp_of, val_of):
    b = ea.Batch()
    expect, singles = [], []
    for k, (f, x) in enumerate(cols_vals):
        col = ea.morph(f, ea.upload_values(x)) if f.kind else ea.upload_values(x)
        cmp, val = cmp_of(k, x), val_of(k, x)
        b.enqueue(col, cmp, val)
        expect.append(rm.masked_sum(x, cmp, val))
        singles.append((col, cmp, val))
    return b, expect, singles


def pick_val(k, x):
    if len(x) == 0:
        return 5
    return int(np.sort(x)[(k * 7919) % len(x)])


def check_results(b, expect):
    res = ea.DeviceU64(max(len(b), 1))
    b.run(res)
    got = res.get().tolist()[:len(b)]
    bad = [(k, g, e) for k, (g, e) in enumerate(zip(got, expect)) if g != e]
    assert not bad, bad[:5]
    return res


def test_batch_edge_columns_all_cmps():
    cv = edge_columns()
    b, expect, singles = make_batch(cv, lambda k, x: CMPS[k % 6], pick_val)
    check_results(b, expect)
    # the single-column path agrees entry by entry
    for (col, cmp, val), e in zip(singles, expect):
        assert int(ea.select_sum(cmp, val, col).get()[0]) == e


@pytest.mark.parametrize("cfg", [(1, 1, 1, 64, 64, 0), (8, 32, 8, 16384, 32768, 6), (3, 3, 3, 100, 1000, 8),
                                 (16, 64, 16, 64, 64, 0), (2, 64, 16, 64, 64, 8)])
def test_batch_configs(cfg):
    cv = edge_columns()
    b, expect, _ = make_batch(cv, lambda k, x: MS_CMP_LT, pick_val)
    b.config = BatchConfig(*cfg)
    check_results(b, expect)


def test_batch_rejects_bad_config_and_layouts_without_decoder():
    b = ea.Batch()
    with pytest.raises(ea.MsError):
        b.config = BatchConfig(0, 1, 1, 64, 64, 0)
    with pytest.raises(ea.MsError):
        b.config = BatchConfig(1, 1, 1, 63, 64, 0)
    with pytest.raises(ea.MsError):
        b.config = BatchConfig(1, 65, 1, 64, 64, 0)
    with pytest.raises(ea.MsError):
        b.config = BatchConfig(1, 1, 1, 64, 64, 7)
    col = ea.morph(group_simple_f(8), ea.upload_values(datagen.uniform_bits(1, 5000, 12)))
    with pytest.raises(ea.MsError):
        b.enqueue(col, MS_CMP_LT, 100)


def test_batch_mixed_workload_tuned():
    cols = list(datagen.mixed_columns(datagen.BASE_SEED + 3, 300, 3.0, 5.0))
    b, expect, _ = make_batch([(f, x) for _, f, x, _ in cols], lambda k, x: MS_CMP_LT,
                              lambda k, x: cols[k][3])
    res = check_results(b, expect)
    cfg = b.tune(res, reps=2)
    assert cfg.wg_per_cu >= 1
    check_results(b, expect)
    ents, items, nbytes = b.stats()
    assert ents == 300 and items >= 300 and nbytes > 0


def test_batch_in_a_plan_replays():
    from engine_amd import hip
    cols = list(datagen.mixed_columns(datagen.BASE_SEED + 4, 60, 3.0, 4.5))
    b, expect, _ = make_batch([(f, x) for _, f, x, _ in cols], lambda k, x: MS_CMP_GE,
                              lambda k, x: cols[k][3])
    st = hip.Stream()
    res = ea.DeviceU64(len(b))
    b.run(res, st.handle)                       # builds the queue
    st.synchronize()
    plan = ea.Plan(1 << 20)
    with plan.record(st.handle):
        b.run(res, st.handle)
    for rep in range(3):
        other = ea.DeviceU64(len(b))
        b.run(other, st.handle)
        plan.launch(st.handle)
        st.synchronize()
        assert res.get().tolist() == expect, rep
        assert other.get().tolist() == expect, rep
    plan.free()


def test_batch_recorded_in_a_plan_is_pinned():
    """A captured plan reads the batch's device queue: while the plan is
    alive, rebuilding (enqueue, set_config, tune) or freeing the batch is
    refused; after the plan is freed it works again.  A batch that never
    ran cannot be recorded (its queue would be built inside the capture)."""
    from engine_amd import hip
    x = datagen.uniform_bits(77, 5000, 20)
    col = ea.morph(dynamic_vbp_f(512, 2048, 8), ea.upload_values(x))
    st = hip.Stream()
    res = ea.DeviceU64(1)
    fresh = ea.Batch()
    fresh.enqueue(col, MS_CMP_LT, 1 << 19)
    p0 = ea.Plan(1 << 20)
    with pytest.raises(ea.MsError) as e:
        with p0.record(st.handle):
            fresh.run(res, st.handle)
    assert e.value.code == 1
    p0.free()
    b = ea.Batch()
    b.enqueue(col, MS_CMP_LT, 1 << 19)
    b.run(res, st.handle)
    st.synchronize()
    plan = ea.Plan(1 << 20)
    with plan.record(st.handle):
        b.run(res, st.handle)
    for call in (lambda: b.enqueue(col, MS_CMP_GE, 0), lambda: b.tune(res, 1, st.handle),
                 lambda: setattr(b, "config", BatchConfig(4, 16, 2, 4096, 8192, 0)), b.free):
        with pytest.raises(ea.MsError) as e:
            call()
        assert e.value.code == 1
    res.zero()
    plan.launch(st.handle)
    st.synchronize()
    assert res.get().tolist() == [rm.masked_sum(x, MS_CMP_LT, 1 << 19)]
    plan.free()
    b.enqueue(col, MS_CMP_GE, 0)
    res2 = ea.DeviceU64(2)
    b.run(res2, st.handle)
    st.synchronize()
    assert res2.get().tolist() == [rm.masked_sum(x, MS_CMP_LT, 1 << 19), rm.masked_sum(x, MS_CMP_GE, 0)]
