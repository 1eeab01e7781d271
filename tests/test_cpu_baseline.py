"""The CPU baseline (oracle/ms_cpu.c: streaming select + sum, scalar and
AVX-512 styles, one or more threads) equals numpy and the oracle's
decompress-then-sum on every layout it streams, every comparison and
ragged sizes."""
import numpy as np
import pytest

from engine_amd import datagen
from engine_amd.formats import (MS_CMP_EQ, MS_CMP_GE, MS_CMP_GT, MS_CMP_LE, MS_CMP_LT, MS_CMP_NE,
                                default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, dynamic_vbp_f, rle_f,
                                static_vbp_f, uncompr_f)
from tests import refmodel as rm


def _dvbp(s, n):
    return np.concatenate([datagen.dvbp_random_bw(s, n - n % 512, lo=1, hi=64),
                           datagen.uniform_bits(s + 1, n % 512, 20)])


LAYOUTS = [
    ("dynamic_vbp<512,2048,8>", dynamic_vbp_f(512, 2048, 8), _dvbp),
    ("dynamic_vbp<1024,64,8>", dynamic_vbp_f(1024, 64, 8), _dvbp),
    ("delta+dvbp", default_delta_dynamic_vbp_f(8), lambda s, n: datagen.sorted_geometric(s, n)),
    ("delta+dvbp wide", default_delta_dynamic_vbp_f(8), lambda s, n: datagen.uniform_bits(s, n, 50)),
    ("for+dvbp", default_for_dynamic_vbp_f(8), lambda s, n: datagen.for_base_offset(s, n)),
    ("rle", rle_f(), lambda s, n: datagen.rle_runs(s, n, mean_len=9)),
    ("uncompr", uncompr_f(), lambda s, n: datagen.uniform_bits(s, n, 40)),
]
CMPS = [MS_CMP_EQ, MS_CMP_NE, MS_CMP_LT, MS_CMP_LE, MS_CMP_GT, MS_CMP_GE]


@pytest.mark.parametrize("name,f,gen", LAYOUTS, ids=[x[0] for x in LAYOUTS])
@pytest.mark.parametrize("n", [0, 511, 4096, 100_003])
def test_cpu_baseline_matches(oracle, name, f, gen, n):
    x = np.ascontiguousarray(gen(datagen.BASE_SEED + 60 + n % 7, n), dtype=np.uint64)
    col = oracle.morph(f, x) if f.kind else oracle.from_values(x)
    simds = [0, 1] if oracle.cpu_has_avx512() else [0]
    for cmp in CMPS:
        for val in ([int(x[len(x) // 2]), int(np.median(x)) if len(x) else 0, 0, 2**64 - 1] if len(x) else [5]):
            want = rm.masked_sum(x, cmp, val)
            assert oracle.select_sum(cmp, val, col) == want
            for simd in simds:
                for threads in (1, 3):
                    got = oracle.cpu_select_sum(cmp, val, col, threads, simd)
                    assert got == want, (name, n, cmp, val, simd, threads)


def test_cpu_baseline_refuses_other_layouts(oracle):
    x = datagen.uniform_bits(1, 4096, 12)
    col = oracle.morph(static_vbp_f(12, 1), x)
    with pytest.raises(RuntimeError):
        oracle.cpu_select_sum(MS_CMP_LT, 5, col, 1, 0)
