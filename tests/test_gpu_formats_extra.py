"""static_vbp_f<vbp_padding_l>, k_wise_ns_f<2> and group_simple_f<E> on the
device, through the C-ABI: compressed byte images and meta data equal the
oracle's, uploaded oracle images decompress to the values, select + sum and
the operators (which read them through decompression) are exact, and
prepare_for_random_access matches."""
import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen
from engine_amd.formats import (MS_CMP_GE, MS_CMP_LT, group_simple_f, k_wise_ns_f, static_vbp_padding_f,
                                uncompr_f)
from tests import refmodel as rm

pytestmark = pytest.mark.gpu

FORMATS = [static_vbp_padding_f(1, 8), static_vbp_padding_f(7, 8), static_vbp_padding_f(21, 8),
           static_vbp_padding_f(33, 4), static_vbp_padding_f(64, 8), k_wise_ns_f(2), group_simple_f(8),
           group_simple_f(4)]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def data_for(f, n, seed):
    if f.kind == 6:
        return datagen.uniform_bits(seed, n, f.bw)
    # widths varying along the column: group_simple changes modes, k_wise_ns lengths
    x = datagen.uniform_bits(seed, n, 40)
    sh = (datagen.uniform_bits(seed + 1, max(n // 512, 1), 6) % np.uint64(41)).repeat(512)[:n]
    if len(sh) < n:
        sh = np.concatenate([sh, np.zeros(n - len(sh), np.uint64)])
    return x >> sh


def same(g, o):
    img, meta = g.image()
    assert meta == o.meta(), (meta, o.meta())
    assert img.tobytes() == o.image().tobytes()


@pytest.mark.parametrize("f", FORMATS, ids=repr)
@pytest.mark.parametrize("n", [0, 1, 333, 8192, 100_000 + 13, 1 << 20])
def test_morph_parity(oracle, f, n):
    x = np.ascontiguousarray(data_for(f, n, 100 + n % 17), dtype=np.uint64)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    same(g, o)
    assert np.array_equal(g.values(), x)
    # the oracle's image uploaded (block index built on the host) decodes
    u = ea.upload_image(f, o.image(), *o.meta())
    assert np.array_equal(u.values(), x)
    if n:
        val = int(np.median(x))
        for c in (g, u):
            assert int(ea.select_sum(MS_CMP_LT, val, c).get()[0]) == rm.masked_sum(x, MS_CMP_LT, val)
        assert int(ea.select_sum_noalloc(MS_CMP_GE, val, g).get()[0]) == rm.masked_sum(x, MS_CMP_GE, val)


@pytest.mark.parametrize("f", [static_vbp_padding_f(7, 8), k_wise_ns_f(2), group_simple_f(8)], ids=repr)
def test_operators_read_them(oracle, f):
    n = 50_000 + 3
    x = np.ascontiguousarray(data_for(f, n, 7), dtype=np.uint64)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    val = int(np.median(x))
    sel = ea.select(MS_CMP_LT, val, uncompr_f(), g, 8)
    osel = oracle.select(MS_CMP_LT, val, uncompr_f(), o, 8)
    same(sel, osel)
    pos = np.array([0, n - 1, n // 2, 17], dtype=np.uint64)
    assert np.array_equal(ea.project(uncompr_f(), g, ea.upload_values(pos)).values(), x[pos.astype(np.int64)])
    assert int(ea.agg_sum(g).values()[0]) == rm.masked_sum(x, MS_CMP_GE, 0)
    # morph between the formats and back
    for h in FORMATS[:2] + [group_simple_f(8)]:
        if h.kind == 6 and int(x.max()).bit_length() > h.bw:
            continue
        assert np.array_equal(ea.morph(h, g).values(), x)


@pytest.mark.parametrize("f", [static_vbp_padding_f(7, 8), k_wise_ns_f(2), group_simple_f(8)], ids=repr)
@pytest.mark.parametrize("n", [5, 4096 + 77])
def test_prepare_for_random_access_extra(oracle, f, n):
    x = np.ascontiguousarray(data_for(f, n, 3), dtype=np.uint64)
    g = ea.morph(f, ea.upload_values(x))
    o = oracle.morph(f, x)
    assert g.prepare_for_random_access() == o.prepare_for_random_access()
    same(g, o)
    assert np.array_equal(g.values(), x)


def test_no_write_iterator_and_no_batch():
    g = ea.morph(group_simple_f(8), ea.upload_values(np.arange(1000, dtype=np.uint64)))
    with pytest.raises(ea.MsError):
        ea.select(MS_CMP_LT, 5, group_simple_f(8), g, 8)
    b = ea.Batch()
    with pytest.raises(ea.MsError):
        b.enqueue(g, MS_CMP_LT, 5)
