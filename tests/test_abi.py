"""The C-ABI library loads and exports every entry point include/*.h
declares (no GPU needed: nothing is called)."""
import ctypes
import glob
import os
import re

import engine_amd as ea

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(ms_[a-z_0-9]+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_entry_points()
    assert {"ms_init", "ms_morph", "ms_select", "ms_project", "ms_agg_sum", "ms_select_sum",
            "ms_intersect_sorted", "ms_merge_sorted", "ms_calc_binary"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ea.LIB_PATH)
    missing = [n for n in sorted(declared_entry_points()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_entry_points() <= set(ea.ABI)


def test_calls_without_init_fail_cleanly():
    lib = ea.lib()
    assert lib.ms_version().startswith(b"morphstore-hip")
    f = ea.FormatDesc(2, 0, 512, 2048, 8)
    out = ctypes.c_uint64()
    assert lib.ms_size_max_byte_any_len(ctypes.byref(f), 1000, ctypes.byref(out)) == 0
    bad = ea.FormatDesc(2, 0, 500, 2048, 8)
    assert lib.ms_size_max_byte_any_len(ctypes.byref(bad), 1000, ctypes.byref(out)) == 1


def test_size_functions_match_oracle(oracle):
    from engine_amd.formats import (default_delta_dynamic_vbp_f, default_for_dynamic_vbp_f, dynamic_vbp_f, rle_f,
                                    static_vbp_f, uncompr_f)
    lib = ea.lib()
    out = ctypes.c_uint64()
    for f in [uncompr_f(), static_vbp_f(12, 1), dynamic_vbp_f(512, 2048, 8), default_delta_dynamic_vbp_f(8),
              default_for_dynamic_vbp_f(8), rle_f()]:
        for n in [0, 1, 511, 512, 100000]:
            assert lib.ms_size_max_byte_any_len(ctypes.byref(f), n, ctypes.byref(out)) == 0
            assert out.value == oracle.size_max_byte_any_len(f, n), (f, n)


def test_col_free_refuses_unknown_handles():
    """ms_col_free frees only handles the library handed out (a second free
    or a foreign pointer is refused, not dereferenced); no GPU needed."""
    lib = ctypes.CDLL(ea.LIB_PATH)
    lib.ms_col_free.argtypes = [ctypes.c_void_p]
    assert lib.ms_col_free(None) == 0
    buf = ctypes.create_string_buffer(512)
    assert lib.ms_col_free(ctypes.cast(buf, ctypes.c_void_p)) == 1   # MS_ERR_INVALID
