"""blockwise_individual_f on the device through the C-ABI: byte images and
meta data equal the oracle's for every selector and for given codes, the
codes read back, oracle images uploaded decode, select + sum, the
operators (through decompression), prepare_for_random_access, binary_io
and the persistent batch kernel over the views."""
import numpy as np
import pytest

import engine_amd as ea
from engine_amd.formats import (MS_BWI_SEL_BEST, MS_BWI_SEL_STATIC, MS_BWI_SEL_WORST, MS_CMP_GE, MS_CMP_LT,
                                blockwise_individual_f, dynamic_vbp_f, uncompr_f)
from tests import bwi_data
from tests import refmodel as rm

pytestmark = pytest.mark.gpu

ALL_CODES = [0, 1, 5, 12, 31, 33, 64, 65, 66, 67, 69, 71]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def same(g, o):
    img, meta = g.image()
    assert meta == o.meta(), (meta, o.meta())
    assert img.tobytes() == o.image().tobytes()


@pytest.mark.parametrize("sel", [MS_BWI_SEL_BEST, MS_BWI_SEL_WORST])
@pytest.mark.parametrize("B,nb,rest", [(1024, 0, 5), (1024, 7, 0), (2048, 12, 333), (3072, 5, 1), (65536, 3, 77)])
def test_bwi_selector_parity(oracle, sel, B, nb, rest):
    x = bwi_data.mixture(31 + nb, nb, B, rest)
    f = blockwise_individual_f(B, sel)
    print(f"bwi selector {sel} B={B} nb={nb} rest={rest}", flush=True)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    same(g, o)
    assert np.array_equal(g.block_codes(), oracle.block_codes(o))
    assert np.array_equal(g.values(), x)
    val = int(np.median(x)) if len(x) else 0
    assert int(ea.select_sum(MS_CMP_LT, val, g).get()[0]) == rm.masked_sum(x, MS_CMP_LT, val)


@pytest.mark.parametrize("B", [1024, 2048, 3072, 16384])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bwi_given_codes_parity(oracle, B, seed):
    nb, rest = 10, 17 * seed
    x = bwi_data.mixture(seed, nb, B, rest)
    codes = bwi_data.fitting_codes(x, B, seed, ALL_CODES)
    f = blockwise_individual_f(B)
    print(f"bwi codes B={B} seed={seed} codes={codes.tolist()}", flush=True)
    g = ea.morph_blockwise(f, ea.upload_values(x), codes)
    o = oracle.morph_codes(f, x, codes)
    same(g, o)
    assert np.array_equal(g.block_codes(), codes)
    assert np.array_equal(g.values(), x)
    for cmp in (MS_CMP_LT, MS_CMP_GE):
        val = int(x[len(x) // 3])
        assert int(ea.select_sum(cmp, val, g).get()[0]) == rm.masked_sum(x, cmp, val)
    # the oracle's image uploaded: views found by walking it
    u = ea.upload_image(f, o.image(), *o.meta())
    assert np.array_equal(u.block_codes(), codes)
    assert np.array_equal(u.values(), x)
    val = int(np.median(x))
    assert int(ea.select_sum(MS_CMP_LT, val, u).get()[0]) == rm.masked_sum(x, MS_CMP_LT, val)
    # a same-format morph copies the views too
    assert np.array_equal(ea.morph(f, u).values(), x)


def test_bwi_static_selector(oracle):
    B = 2048
    x = bwi_data.mixture(4, 6, B, 9, kinds=2)
    for code in (65, 67, 69, 71, 64, 0):
        f = blockwise_individual_f(B, MS_BWI_SEL_STATIC, code)
        g = ea.morph(f, ea.upload_values(x))
        same(g, oracle.morph(f, x))
        assert g.block_codes().tolist() == [code] * 6


def test_bwi_refused():
    x = ea.upload_values(np.arange(4096, dtype=np.uint64))
    for bad in (68, 70, 72):
        with pytest.raises(ea.MsError):
            ea.morph_blockwise(blockwise_individual_f(1024), x, np.full(4, bad, np.uint8))
    with pytest.raises(ea.MsError):
        ea.morph_blockwise(blockwise_individual_f(1024), x, np.full(3, 65, np.uint8))   # wrong count
    with pytest.raises(ea.MsError):
        ea.morph(blockwise_individual_f(1000), x)                                      # not a multiple of 1024


def test_bwi_operators(oracle):
    B = 2048
    x = bwi_data.mixture(9, 8, B, 100)
    f = blockwise_individual_f(B)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    val = int(np.median(x))
    same(ea.select(MS_CMP_LT, val, uncompr_f(), g, 8), oracle.select(MS_CMP_LT, val, uncompr_f(), o, 8))
    pos = np.array([0, len(x) - 1, 5000, 17], dtype=np.uint64)
    assert np.array_equal(ea.project(uncompr_f(), g, ea.upload_values(pos)).values(), x[pos.astype(np.int64)])
    assert int(ea.agg_sum(g).values()[0]) == rm.masked_sum(x, MS_CMP_GE, 0)
    d = dynamic_vbp_f(512, 64, 8)
    same(ea.morph(d, g), oracle.morph(d, x))
    with pytest.raises(ea.MsError):
        ea.select(MS_CMP_LT, val, f, g, 8)   # no write-iterator for this format


@pytest.mark.parametrize("B,rest", [(1024, 300), (2048, 1), (4096, 4000)])
def test_bwi_prepare_and_binary_io(oracle, tmp_path, B, rest):
    x = bwi_data.mixture(B + rest, 4, B, rest)
    f = blockwise_individual_f(B)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    assert g.prepare_for_random_access() == o.prepare_for_random_access()
    same(g, o)
    assert np.array_equal(g.values(), x)
    path = str(tmp_path / "bwi.col")
    g.store(path)
    h = ea.load(f, path)
    assert h.image()[1] == g.image()[1]
    assert np.array_equal(h.values(), x)
    assert np.array_equal(h.block_codes(), g.block_codes())


def test_bwi_batch():
    """Views of every batchable code as items of the persistent kernel."""
    cols, xs = [], []
    for i, B in enumerate([1024, 2048, 3072, 65536]):
        x = bwi_data.mixture(50 + i, 5, B, 11 * i)
        codes = bwi_data.fitting_codes(x, B, i, [0, 3, 40, 64, 65, 67, 69, 71])
        cols.append(ea.morph_blockwise(blockwise_individual_f(B), ea.upload_values(x), codes))
        xs.append(x)
    b = ea.Batch()
    vals = []
    for c, x in zip(cols, xs):
        v = int(np.median(x))
        vals.append(v)
        b.enqueue(c, MS_CMP_LT, v)
    res = ea.DeviceU64(len(cols))
    b.run(res)
    got = res.get()
    for k, (x, v) in enumerate(zip(xs, vals)):
        assert int(got[k]) == rm.masked_sum(x, MS_CMP_LT, v), k
    gs = ea.morph_blockwise(blockwise_individual_f(1024), ea.upload_values(xs[0][:2048]), np.array([66, 65], np.uint8))
    with pytest.raises(ea.MsError):
        b.enqueue(gs, MS_CMP_LT, 5)
