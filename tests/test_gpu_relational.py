"""group, join, semi join and binary_io on the device, through the C-ABI:
the reference's known-answer tests, and equality with the oracle's scalar
restatements (hash-map group ids in order of first appearance, nested-loop
joins) on compressed and uncompressed inputs; binary_io files byte for
byte as the reference writes them."""
import json
import os
import struct

import numpy as np
import pytest

import engine_amd as ea
from engine_amd import datagen
from engine_amd.formats import (MS_ERR_SIZE, default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f,
                                dynamic_vbp_f, rle_f, static_vbp_f, uncompr_f)

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "reference_kat.json")))
U = uncompr_f()


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ea.init(0)


def col(v, f=None):
    c = ea.upload_values(np.asarray(v, dtype=np.uint64))
    return ea.morph(f, c) if f is not None and f.kind else c


def test_kat_group_join():
    k = KAT["group_unary"]
    g, e = ea.group(U, U, col(k["data"]))
    assert g.values().tolist() == k["gr"] and e.values().tolist() == k["ext"]
    k = KAT["group_binary"]
    g, e = ea.group(U, U, col(k["data"]), col(k["in_gr"]))
    assert g.values().tolist() == k["gr"] and e.values().tolist() == k["ext"]
    k = KAT["join"]
    a, b = ea.join(U, U, col(k["l"]), col(k["r"]))
    assert a.values().tolist() == k["out_l"] and b.values().tolist() == k["out_r"]
    k = KAT["semi_join"]
    assert ea.semi_join(U, col(k["l"]), col(k["r"])).values().tolist() == k["out"]


@pytest.mark.parametrize("in_f", [U, dynamic_vbp_f(512, 2048, 8), default_for_dynamic_vbp_f(8), rle_f()],
                         ids=repr)
@pytest.mark.parametrize("n", [0, 1, 5000, 300_000 + 7])
def test_group_matches_oracle(oracle, in_f, n):
    x = datagen.uniform_bits(datagen.BASE_SEED + 70 + n % 13, n, 11)
    if in_f.kind == 5:
        x = datagen.rle_runs(datagen.BASE_SEED + 71, n, mean_len=5, value_bits=8)
    y = datagen.uniform_bits(datagen.BASE_SEED + 72, n, 3)
    for gin in (None, y):
        og, oe = oracle.group(oracle.from_values(x), oracle.from_values(gin) if gin is not None else None)
        for gf, ef in [(U, U), (default_delta_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8))]:
            g, e = ea.group(gf, ef, col(x, in_f), col(gin, dynamic_vbp_f(512, 2048, 8)) if gin is not None else None)
            assert np.array_equal(g.values(), og.values()), (in_f, n, gin is None, gf)
            assert np.array_equal(e.values(), oe.values())
            if gf.kind == 0 and n:
                assert g.image()[0].tobytes() == og.image().tobytes()


def test_group_binary_size_mismatch():
    with pytest.raises(ea.MsError) as ex:
        ea.group(U, U, col([1, 2, 3]), col([0, 0]))
    assert ex.value.code == MS_ERR_SIZE


@pytest.mark.parametrize("na,nb", [(0, 10), (10, 0), (50_000, 3_000), (3_000, 50_000), (4096, 4096)])
def test_join_matches_oracle(oracle, na, nb):
    a = datagen.uniform_bits(datagen.BASE_SEED + 80 + na, na, 12)
    b = datagen.uniform_bits(datagen.BASE_SEED + 81 + nb, nb, 12)
    ol, orr = oracle.join(oracle.from_values(a), oracle.from_values(b))
    for lf, rf, inf in [(U, U, U), (default_delta_dynamic_vbp_f(8), dynamic_vbp_f(512, 2048, 8),
                                    dynamic_vbp_f(512, 2048, 8))]:
        l, r = ea.join(lf, rf, col(a, inf), col(b, inf))
        assert np.array_equal(l.values(), ol.values()) and np.array_equal(r.values(), orr.values()), (na, nb, lf)
    s = ea.semi_join(default_delta_dynamic_vbp_f(8), col(a, static_vbp_f(12, 8)), col(b))
    assert np.array_equal(s.values(), oracle.semi_join(oracle.from_values(a), oracle.from_values(b)).values())


BIO_FORMATS = [U, static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8), default_delta_dynamic_vbp_f(8),
               default_for_dynamic_vbp_f(8), rle_f()]


@pytest.mark.parametrize("f", BIO_FORMATS, ids=repr)
@pytest.mark.parametrize("n", [0, 3, 4096, 100_000 + 5])
def test_binary_io_round_trip(oracle, tmp_path, f, n):
    """binary_io<F>::store writes count, sizeUsed and the image (the gap as
    zeros); load of that file gives back the same column."""
    x = datagen.uniform_bits(datagen.BASE_SEED + 90, n, 12 if f.kind == 1 else 30)
    if f.kind == 3:
        x = np.sort(x)
    if f.kind == 5:
        x = datagen.rle_runs(datagen.BASE_SEED + 91, n, mean_len=7)
    g = col(x, f)
    o = oracle.morph(f, x) if f.kind else oracle.from_values(x)
    path = tmp_path / "col.bin"
    g.store(str(path))
    raw = path.read_bytes()
    assert raw[:16] == struct.pack("<QQ", n, o.size_used)
    assert raw[16:] == o.image().tobytes()
    h = ea.load(f, str(path))
    img, meta = h.image()
    assert meta == o.meta() and img.tobytes() == o.image().tobytes()
    assert np.array_equal(h.values(), x)


def test_binary_io_errors(tmp_path):
    with pytest.raises(ea.MsError):
        ea.load(U, str(tmp_path / "missing.bin"))
    p = tmp_path / "short.bin"
    p.write_bytes(struct.pack("<QQ", 10, 80) + b"\0" * 40)
    with pytest.raises(ea.MsError):
        ea.load(U, str(p))
    # an image whose size does not fit a morph of `count` values
    p.write_bytes(struct.pack("<QQ", 1000, 64) + b"\1" * 64)
    with pytest.raises(ea.MsError):
        ea.load(dynamic_vbp_f(512, 2048, 8), str(p))
