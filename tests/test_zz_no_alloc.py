"""ms_select_sum_noalloc (the fused select + agg_sum that allocates no
device memory) through the C-ABI: its sums equal ms_select_sum's on every
layout, on two calls in a row with the second on a smaller input (the
library's per-device temporaries keep the first call's contents), and a
column beyond the temporaries' capacity is refused."""
import ctypes

import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_GE, MS_CMP_LT, MS_CMP_NE, MS_ERR_CAPACITY,
                                default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, delta_f, dynamic_vbp_f,
                                rle_f, static_vbp_f, uncompr_f)
from tests import refmodel as rm

pytestmark = pytest.mark.gpu

def _dvbp(s, n):
    # whole 512-value blocks of random widths, then a ragged rest
    return np.concatenate([datagen.dvbp_random_bw(s, n - n % 512), datagen.uniform_bits(s + 1, n % 512, 20)])


CASES = [
    ("uncompr", uncompr_f(), lambda s, n: datagen.uniform_bits(s, n, 30)),
    ("static_vbp<12,1>", static_vbp_f(12, 1), lambda s, n: datagen.uniform_bits(s, n, 12)),
    ("static_vbp<7,8>", static_vbp_f(7, 8), lambda s, n: datagen.uniform_bits(s, n, 7)),
    ("dynamic_vbp<512,2048,8>", dynamic_vbp_f(512, 2048, 8), _dvbp),
    ("delta+dvbp", default_delta_dynamic_vbp_f(8), lambda s, n: datagen.sorted_geometric(s, n)),
    ("delta+static_vbp<40,8>", delta_f(1024, 8, static_vbp_f(40, 8)), lambda s, n: datagen.sorted_geometric(s, n)),
    ("for+dvbp", default_for_dynamic_vbp_f(8), lambda s, n: datagen.for_base_offset(s, n)),
    ("rle", rle_f(), lambda s, n: datagen.rle_runs(s, n)),
]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def _sum_both(col, cmp, val):
    a = ea.select_sum(cmp, val, col).get()[0]
    b = ea.select_sum_noalloc(cmp, val, col).get()[0]
    return int(a), int(b)


@pytest.mark.parametrize("name,f,gen", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_select_sum_twice(name, f, gen):
    # the first call fills the shared temporaries; the second, smaller one
    # must not pick up anything the first left behind
    for k, n in enumerate([(1 << 20) - 333, 70_001]):
        x = np.ascontiguousarray(gen(0x4D6F7270 + 40 + k, n), dtype=np.uint64)
        col = ea.morph(f, ea.upload_values(x))
        for cmp, val in [(MS_CMP_LT, int(np.median(x))), (MS_CMP_GE, 0), (MS_CMP_EQ, int(x[n // 2])),
                         (MS_CMP_NE, int(x[n // 3]))]:
            got, want_abi = _sum_both(col, cmp, val)
            assert got == want_abi == rm.masked_sum(x, cmp, val), (name, n, cmp)


def test_noalloc_refuses_beyond_its_temporaries():
    n = (1 << 20) + 4096
    x = datagen.sorted_geometric(0x4D6F7270 + 50, n)
    col = ea.morph(delta_f(1024, 8, static_vbp_f(40, 8)), ea.upload_values(x))
    r = ea.DeviceU64()
    rc = ea.lib().ms_select_sum_noalloc(r.p, MS_CMP_LT, 1 << 30, col.h, None)
    assert rc == MS_ERR_CAPACITY, rc
    assert b"temporary" in ea.lib().ms_last_error()
    # the allocating entry point takes the same column
    assert int(ea.select_sum(MS_CMP_LT, 1 << 30, col).get()[0]) == rm.masked_sum(x, MS_CMP_LT, 1 << 30)


def test_noalloc_null_arguments():
    col = ea.upload_values(np.arange(10, dtype=np.uint64))
    assert ea.lib().ms_select_sum_noalloc(None, MS_CMP_LT, 5, col.h, None) != 0
    r = ea.DeviceU64()
    assert ea.lib().ms_select_sum_noalloc(r.p, 99, 5, col.h, None) != 0
    assert ea.lib().ms_select_sum_noalloc(r.p, MS_CMP_LT, 5, ctypes.c_void_p(), None) != 0
