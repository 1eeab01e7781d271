"""ctypes binding of the CPU oracle (oracle/ms_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/,
by __graft_entry__.smoke() and by bench.py's CPU-baseline leg; the product
(engine_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

from engine_amd.formats import FormatDesc

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmsoracle.so")


class OcCol(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_uint8)), ("count", ctypes.c_uint64),
                ("size_used", ctypes.c_uint64), ("size_compr", ctypes.c_uint64),
                ("size_alloc", ctypes.c_uint64), ("f", FormatDesc)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P, U64, U32 = ctypes.POINTER, ctypes.c_uint64, ctypes.c_uint32
        FD, C = P(FormatDesc), P(OcCol)
        sig = {
            "oc_effective_bitwidth": (ctypes.c_uint, [U64]),
            "oc_size_max_byte": (ctypes.c_int, [FD, U64, P(U64)]),
            "oc_size_max_byte_any_len": (ctypes.c_int, [FD, U64, P(U64)]),
            "oc_block_size": (ctypes.c_int, [FD, P(U64)]),
            "oc_rest_offset": (ctypes.c_int, [FD, U64, P(U64)]),
            "oc_morph_batch_compress": (ctypes.c_int, [FD, ctypes.c_void_p, U64, ctypes.c_void_p, P(U64)]),
            "oc_morph_batch_decompress": (ctypes.c_int, [FD, ctypes.c_void_p, U64, ctypes.c_void_p, P(U64)]),
            "oc_col_from_values": (ctypes.c_int, [C, ctypes.c_void_p, U64]),
            "oc_morph_compress": (ctypes.c_int, [C, FD, ctypes.c_void_p, U64]),
            "oc_col_from_image": (ctypes.c_int, [C, FD, ctypes.c_void_p, U64, U64, U64]),
            "oc_morph_decompress": (ctypes.c_int, [C, ctypes.c_void_p]),
            "oc_count_values_uncompr": (U64, [C]),
            "oc_col_free": (None, [C]),
            "oc_wit_emit": (ctypes.c_int, [C, FD, ctypes.c_void_p, U64, U64, P(U64)]),
            "oc_select": (ctypes.c_int, [C, U32, U64, FD, C, U32]),
            "oc_project": (ctypes.c_int, [C, FD, C, C, U32]),
            "oc_between": (ctypes.c_int, [C, U32, U64, U32, U64, FD, C, U32]),
            "oc_group": (ctypes.c_int, [C, C, C, C]),
            "oc_join": (ctypes.c_int, [C, C, C, C]),
            "oc_semi_join": (ctypes.c_int, [C, C, C]),
            "oc_agg_sum": (ctypes.c_int, [C, C]),
            "oc_intersect_sorted": (ctypes.c_int, [C, FD, C, C]),
            "oc_merge_sorted": (ctypes.c_int, [C, FD, C, C]),
            "oc_calc_binary": (ctypes.c_int, [C, U32, FD, C, C]),
            "oc_select_sum": (ctypes.c_int, [P(U64), U32, U64, C]),
            "oc_select_sum_stream": (ctypes.c_int, [P(U64), U32, U64, C, ctypes.c_int]),
            "oc_cpu_has_avx512": (ctypes.c_int, []),
            "oc_prepare_for_random_access": (ctypes.c_int, [C, P(ctypes.c_int)]),
            "oc_cpu_select_sum": (ctypes.c_int, [P(U64), U32, U64, C, ctypes.c_int, ctypes.c_int]),
            "oc_bwi_inner": (ctypes.c_int, [ctypes.c_uint, FD]),
            "oc_morph_compress_codes": (ctypes.c_int, [C, FD, ctypes.c_void_p, U64, ctypes.c_void_p]),
            "oc_bwi_codes": (ctypes.c_int, [C, ctypes.c_void_p, U64, P(U64)]),
            "oc_bwi_choose": (ctypes.c_int, [FD, ctypes.c_void_p, P(ctypes.c_uint)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


class Column:
    """A host column of the oracle: byte image + meta data."""

    def __init__(self):
        self.c = OcCol()

    def __del__(self):
        if _lib is not None and self.c.data:
            _lib.oc_col_free(ctypes.byref(self.c))

    @property
    def count(self):
        return self.c.count

    @property
    def size_used(self):
        return self.c.size_used

    @property
    def size_compr(self):
        return self.c.size_compr

    @property
    def size_alloc(self):
        return self.c.size_alloc

    @property
    def fmt(self):
        return self.c.f

    def meta(self):
        return (self.c.count, self.c.size_used, self.c.size_compr)

    def image(self):
        """Bytes [0, size_used) as a uint8 array (gap bytes are zero)."""
        n = self.c.size_used
        if n == 0:
            return np.zeros(0, np.uint8)
        return np.ctypeslib.as_array(self.c.data, shape=(n,)).copy()

    def values(self):
        out = np.zeros(self.c.count + 1, np.uint64)
        _check(lib().oc_morph_decompress(ctypes.byref(self.c), out.ctypes.data), "decompress")
        return out[:self.c.count]

    def count_uncompr(self):
        return lib().oc_count_values_uncompr(ctypes.byref(self.c))

    def prepare_for_random_access(self):
        """column<F>::prepare_for_random_access; True if the column changed."""
        acted = ctypes.c_int()
        _check(lib().oc_prepare_for_random_access(ctypes.byref(self.c), ctypes.byref(acted)),
               "prepare_for_random_access")
        return bool(acted.value)

    def raw(self, nbytes):
        """The first nbytes of the buffer (past sizeUsed too)."""
        return np.ctypeslib.as_array(self.c.data, shape=(nbytes,)).copy()


def _u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def from_values(v):
    v = _u64(v)
    c = Column()
    _check(lib().oc_col_from_values(ctypes.byref(c.c), v.ctypes.data, len(v)), "from_values")
    return c


def from_image(f, image, count, size_used, size_compr):
    image = np.ascontiguousarray(image, dtype=np.uint8)
    c = Column()
    _check(lib().oc_col_from_image(ctypes.byref(c.c), ctypes.byref(f), image.ctypes.data, count, size_used,
                                   size_compr), "from_image")
    return c


def morph(f, v):
    """morph<ve, f, uncompr_f>(column of v)."""
    v = _u64(v)
    c = Column()
    _check(lib().oc_morph_compress(ctypes.byref(c.c), ctypes.byref(f), v.ctypes.data, len(v)), "morph")
    return c


def morph_codes(f, v, codes):
    """morph<ve, blockwise_individual_f, uncompr_f> with the block codes given."""
    v = _u64(v)
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    c = Column()
    _check(lib().oc_morph_compress_codes(ctypes.byref(c.c), ctypes.byref(f), v.ctypes.data, len(v),
                                         codes.ctypes.data), "morph_codes")
    return c


def block_codes(col):
    """The format code of every block of a blockwise_individual_f column."""
    n = ctypes.c_uint64()
    _check(lib().oc_bwi_codes(ctypes.byref(col.c), None, 0, ctypes.byref(n)), "block_codes")
    out = np.zeros(max(n.value, 1), np.uint8)
    _check(lib().oc_bwi_codes(ctypes.byref(col.c), out.ctypes.data, n.value, ctypes.byref(n)), "block_codes")
    return out[:n.value]


def bwi_inner(code):
    """The default format of a format code (None if it has none)."""
    d = FormatDesc()
    return d if lib().oc_bwi_inner(code, ctypes.byref(d)) == 0 else None


def bwi_choose(f, block):
    """The format selector of blockwise_individual_f f over one block."""
    block = _u64(block)
    code = ctypes.c_uint()
    _check(lib().oc_bwi_choose(ctypes.byref(f), block.ctypes.data, ctypes.byref(code)), "bwi_choose")
    return code.value


def morph_batch_compress(f, v):
    v = _u64(v)
    cap = ctypes.c_uint64()
    bs = f.block_size() or 1
    _check(lib().oc_size_max_byte(ctypes.byref(f), max(len(v), bs), ctypes.byref(cap)), "size")
    out = np.zeros(cap.value + 64 + 16 * len(v), np.uint8)
    nbytes = ctypes.c_uint64()
    _check(lib().oc_morph_batch_compress(ctypes.byref(f), v.ctypes.data, len(v), out.ctypes.data,
                                         ctypes.byref(nbytes)), "morph_batch")
    return out[:nbytes.value].copy()


def morph_batch_decompress(f, image, n):
    image = np.ascontiguousarray(image, dtype=np.uint8)
    out = np.zeros(n + 1, np.uint64)
    used = ctypes.c_uint64()
    _check(lib().oc_morph_batch_decompress(ctypes.byref(f), image.ctypes.data, n, out.ctypes.data,
                                           ctypes.byref(used)), "morph_batch_decompress")
    return out[:n], used.value


def wit_emit(f, v):
    v = _u64(v)
    c = Column()
    app = ctypes.c_uint64()
    _check(lib().oc_wit_emit(ctypes.byref(c.c), ctypes.byref(f), v.ctypes.data, len(v), 0,
                             ctypes.byref(app)), "wit_emit")
    return c, app.value


def select(cmp, val, out_f, col, lanes):
    c = Column()
    _check(lib().oc_select(ctypes.byref(c.c), cmp, val, ctypes.byref(out_f), ctypes.byref(col.c), lanes),
           "select")
    return c


def between(cmp_lo, val_lo, cmp_hi, val_hi, out_f, col, lanes):
    c = Column()
    _check(lib().oc_between(ctypes.byref(c.c), cmp_lo, val_lo, cmp_hi, val_hi, ctypes.byref(out_f),
                            ctypes.byref(col.c), lanes), "between")
    return c


def group(data, in_gr=None):
    """(group ids, extents) as uncompressed columns."""
    g, e = Column(), Column()
    _check(lib().oc_group(ctypes.byref(g.c), ctypes.byref(e.c), ctypes.byref(in_gr.c) if in_gr else None,
                          ctypes.byref(data.c)), "group")
    return g, e


def join(l, r):
    a, b = Column(), Column()
    _check(lib().oc_join(ctypes.byref(a.c), ctypes.byref(b.c), ctypes.byref(l.c), ctypes.byref(r.c)), "join")
    return a, b


def semi_join(l, r):
    c = Column()
    _check(lib().oc_semi_join(ctypes.byref(c.c), ctypes.byref(l.c), ctypes.byref(r.c)), "semi_join")
    return c


def project(out_f, data, pos, lanes):
    c = Column()
    _check(lib().oc_project(ctypes.byref(c.c), ctypes.byref(out_f), ctypes.byref(data.c),
                            ctypes.byref(pos.c), lanes), "project")
    return c


def agg_sum(col):
    c = Column()
    _check(lib().oc_agg_sum(ctypes.byref(c.c), ctypes.byref(col.c)), "agg_sum")
    return c


def intersect_sorted(out_f, l, r):
    c = Column()
    _check(lib().oc_intersect_sorted(ctypes.byref(c.c), ctypes.byref(out_f), ctypes.byref(l.c),
                                     ctypes.byref(r.c)), "intersect_sorted")
    return c


def merge_sorted(out_f, l, r):
    c = Column()
    _check(lib().oc_merge_sorted(ctypes.byref(c.c), ctypes.byref(out_f), ctypes.byref(l.c),
                                 ctypes.byref(r.c)), "merge_sorted")
    return c


def calc_binary(op, out_f, l, r):
    c = Column()
    rc = lib().oc_calc_binary(ctypes.byref(c.c), op, ctypes.byref(out_f), ctypes.byref(l.c), ctypes.byref(r.c))
    if rc:
        return rc
    return c


def select_sum(cmp, val, col):
    r = ctypes.c_uint64()
    _check(lib().oc_select_sum(ctypes.byref(r), cmp, val, ctypes.byref(col.c)), "select_sum")
    return r.value


def select_sum_stream(cmp, val, col, threads=1):
    r = ctypes.c_uint64()
    _check(lib().oc_select_sum_stream(ctypes.byref(r), cmp, val, ctypes.byref(col.c), threads),
           "select_sum_stream")
    return r.value


def cpu_has_avx512():
    return bool(lib().oc_cpu_has_avx512())


def cpu_select_sum(cmp, val, col, threads=1, simd=0):
    """The CPU baseline's streaming select + sum (simd 0 scalar, 1 AVX-512)."""
    r = ctypes.c_uint64()
    _check(lib().oc_cpu_select_sum(ctypes.byref(r), cmp, val, ctypes.byref(col.c), threads, simd),
           "cpu_select_sum")
    return r.value


def size_max_byte_any_len(f, n):
    r = ctypes.c_uint64()
    _check(lib().oc_size_max_byte_any_len(ctypes.byref(f), n, ctypes.byref(r)), "size_any_len")
    return r.value


def rest_offset(f, compr_end):
    r = ctypes.c_uint64()
    _check(lib().oc_rest_offset(ctypes.byref(f), compr_end, ctypes.byref(r)), "rest_offset")
    return r.value
