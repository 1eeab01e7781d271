"""engine_amd -- MorphStore's compression-aware execution path on MI355X.

A thin Python binding of the C-ABI in include/morphstore_hip.h (the HIP
kernels live in engine_amd/csrc, built into engine_amd/libmorphstore_hip.so).
There is no CPU fallback: importing the library fails loudly when the shared
object is missing, and every call needs a gfx950 device.
"""
import ctypes
import os

import numpy as np

from .formats import BatchConfig, ColMeta, FormatDesc  # noqa: F401
from .hip import Stream

_HERE = os.path.dirname(os.path.abspath(__file__))
# MS_LIB_PATH selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("MS_LIB_PATH") or os.path.join(_HERE, "libmorphstore_hip.so")

_lib = None
_inited = False

# every entry point the header declares: name -> (restype, argtypes)
P, U64, U32, VP = ctypes.POINTER, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
_FD, _CM = P(FormatDesc), P(ColMeta)
_COLP = P(VP)
ABI = {
    "ms_init": (ctypes.c_int, [ctypes.c_int]),
    "ms_last_error": (ctypes.c_char_p, []),
    "ms_version": (ctypes.c_char_p, []),
    "ms_size_max_byte_any_len": (ctypes.c_int, [_FD, U64, P(U64)]),
    "ms_size_max_byte": (ctypes.c_int, [_FD, U64, P(U64)]),
    "ms_block_size": (ctypes.c_int, [_FD, P(U64)]),
    "ms_col_upload": (ctypes.c_int, [_COLP, _FD, VP, _CM, VP]),
    "ms_col_upload_values": (ctypes.c_int, [_COLP, VP, U64, VP]),
    "ms_col_from_device": (ctypes.c_int, [_COLP, VP, U64, VP]),
    "ms_col_get_meta": (ctypes.c_int, [VP, _CM, VP]),
    "ms_col_download": (ctypes.c_int, [VP, VP, U64, _CM, VP]),
    "ms_col_format": (ctypes.c_int, [VP, _FD]),
    "ms_col_device_data": (VP, [VP]),
    "ms_col_count_bound": (U64, [VP]),
    "ms_col_free": (ctypes.c_int, [VP]),
    "ms_col_check": (ctypes.c_int, [VP, VP]),
    "ms_col_store": (ctypes.c_int, [VP, ctypes.c_char_p, VP]),
    "ms_col_load": (ctypes.c_int, [_COLP, _FD, ctypes.c_char_p, VP]),
    "ms_col_prepare_for_random_access": (ctypes.c_int, [VP, P(ctypes.c_int), VP]),
    "ms_morph": (ctypes.c_int, [_COLP, _FD, VP, VP]),
    "ms_morph_blockwise": (ctypes.c_int, [_COLP, _FD, VP, VP, U64, VP]),
    "ms_col_block_codes": (ctypes.c_int, [VP, VP, U64, P(U64)]),
    "ms_batch_view_times": (ctypes.c_int, [VP, VP, U32, VP, U64, P(U64), VP]),
    "ms_select": (ctypes.c_int, [_COLP, U32, U64, _FD, VP, U32, VP]),
    "ms_between": (ctypes.c_int, [_COLP, U32, U64, U32, U64, _FD, VP, U32, VP]),
    "ms_project": (ctypes.c_int, [_COLP, _FD, VP, VP, U32, VP]),
    "ms_agg_sum": (ctypes.c_int, [_COLP, VP, VP]),
    "ms_select_sum": (ctypes.c_int, [VP, U32, U64, VP, VP]),
    "ms_select_sum_noalloc": (ctypes.c_int, [VP, U32, U64, VP, VP]),
    "ms_intersect_sorted": (ctypes.c_int, [_COLP, _FD, VP, VP, VP]),
    "ms_merge_sorted": (ctypes.c_int, [_COLP, _FD, VP, VP, VP]),
    "ms_calc_binary": (ctypes.c_int, [_COLP, U32, _FD, VP, VP, VP]),
    "ms_group": (ctypes.c_int, [_COLP, _COLP, _FD, _FD, VP, VP, VP]),
    "ms_join": (ctypes.c_int, [_COLP, _COLP, _FD, _FD, VP, VP, VP]),
    "ms_semi_join": (ctypes.c_int, [_COLP, _FD, VP, VP, VP]),
    "ms_plan_create": (ctypes.c_int, [_COLP, U64]),
    "ms_plan_begin": (ctypes.c_int, [VP, VP]),
    "ms_plan_end": (ctypes.c_int, [VP, VP]),
    "ms_plan_launch": (ctypes.c_int, [VP, VP]),
    "ms_plan_arena_bytes": (U64, [VP]),
    "ms_plan_free": (ctypes.c_int, [VP]),
    "ms_batch_create": (ctypes.c_int, [_COLP]),
    "ms_batch_enqueue": (ctypes.c_int, [VP, VP, U32, U64, P(U64)]),
    "ms_batch_run": (ctypes.c_int, [VP, VP, VP]),
    "ms_batch_tune": (ctypes.c_int, [VP, VP, U32, P(BatchConfig), VP]),
    "ms_batch_get_config": (ctypes.c_int, [VP, P(BatchConfig)]),
    "ms_batch_set_config": (ctypes.c_int, [VP, P(BatchConfig)]),
    "ms_batch_stats": (ctypes.c_int, [VP, P(U64), P(U64), P(U64)]),
    "ms_batch_free": (ctypes.c_int, [VP]),
    "ms_batch_timing": (ctypes.c_int, [VP, U32]),
    "ms_batch_kernel_times": (ctypes.c_int, [VP, VP, P(U32)]),
    "ms_gen_chunks": (ctypes.c_int, [VP, VP, U64, VP, U32, VP]),
    "ms_chunk_stats": (ctypes.c_int, [VP, U64, VP, VP, VP, VP]),
    "ms_chunk_features": (ctypes.c_int, [VP, U64, VP, VP, VP]),
    "ms_candidate_format": (ctypes.c_int, [U32, U32, _FD]),
    "ms_mlp_create": (ctypes.c_int, [_COLP, U64]),
    "ms_mlp_get_params": (ctypes.c_int, [VP, VP]),
    "ms_mlp_set_params": (ctypes.c_int, [VP, VP]),
    "ms_mlp_loss_grad": (ctypes.c_int, [VP, VP, VP, VP, U64, VP, VP, VP]),
    "ms_mlp_train": (ctypes.c_int, [VP, VP, VP, VP, U64, U32, U32, ctypes.c_float, U64, VP, VP]),
    "ms_mlp_predict": (ctypes.c_int, [VP, VP, U64, VP, VP]),
    "ms_mlp_free": (ctypes.c_int, [VP]),
    "ms_select_formats": (ctypes.c_int, [VP, VP, U64, VP, VP, VP]),
}


class MsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m engine_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(L, name, None)
            if fn is None:   # a stale build: the call fails loudly when made
                continue
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise MsError(rc, lib().ms_last_error().decode())


def init(device=0):
    global _inited
    if not _inited:
        check(lib().ms_init(device))
        _inited = True


class Column:
    """A device column (ms_col*).  Freed when garbage collected.

    `stream` is the stream the column was created on: ms_col_free releases
    the column in order on it, so a Stream object passed here is kept alive
    as long as the column."""

    def __init__(self, handle, stream=None):
        self.h = handle
        self._stream = stream if isinstance(stream, Stream) else None
        if self._stream is not None:
            self._stream._columns += 1

    def __del__(self):
        if self.h and _lib is not None:
            _lib.ms_col_free(self.h)
            self.h = None
        st, self._stream = self._stream, None
        if st is not None:
            st._column_released()

    def meta(self, stream=None):
        m = ColMeta()
        check(lib().ms_col_get_meta(self.h, ctypes.byref(m), stream))
        return m

    @property
    def fmt(self):
        f = FormatDesc()
        check(lib().ms_col_format(self.h, ctypes.byref(f)))
        return f

    def image(self, stream=None):
        m = self.meta(stream)
        buf = np.zeros(max(m.size_used, 1), np.uint8)
        check(lib().ms_col_download(self.h, buf.ctypes.data, len(buf), None, stream))
        return buf[:m.size_used], (m.count, m.size_used, m.size_compr)

    def values(self, stream=None):
        """Decompressed values (through ms_morph to uncompr_f)."""
        u = morph(FormatDesc(0), self, stream)
        img, meta = u.image(stream)
        return np.frombuffer(img.tobytes(), np.uint64)[:meta[0]].copy()

    def check(self, stream=None):
        check(lib().ms_col_check(self.h, stream))

    def store(self, path, stream=None):
        """binary_io<F>::store."""
        check(lib().ms_col_store(self.h, os.fsencode(path), stream))

    def prepare_for_random_access(self, stream=None):
        """column<F>::prepare_for_random_access; True if the column changed."""
        acted = ctypes.c_int()
        check(lib().ms_col_prepare_for_random_access(self.h, ctypes.byref(acted), stream))
        return bool(acted.value)

    def block_codes(self):
        """blockwise_individual_f: the format code of every block."""
        n = ctypes.c_uint64()
        check(lib().ms_col_block_codes(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.uint8)
        check(lib().ms_col_block_codes(self.h, out.ctypes.data, n.value, ctypes.byref(n)))
        return out[:n.value]

    @property
    def device_ptr(self):
        return lib().ms_col_device_data(self.h)


def _new(fn, *args):
    h = VP()
    check(fn(ctypes.byref(h), *args))
    return Column(h, next((a for a in args if isinstance(a, Stream)), None))


def upload_values(values, stream=None):
    init()
    v = np.ascontiguousarray(values, dtype=np.uint64)
    return _new(lib().ms_col_upload_values, v.ctypes.data, len(v), stream)


def from_device(d_values, count, stream=None):
    """An uncompr_f column copied from `count` device values (a pointer)."""
    init()
    return _new(lib().ms_col_from_device, d_values, count, stream)


def upload_image(fmt, image, count, size_used, size_compr, stream=None):
    init()
    img = np.ascontiguousarray(image, dtype=np.uint8)
    m = ColMeta(count, size_used, size_compr, max(size_used, 1))
    return _new(lib().ms_col_upload, ctypes.byref(fmt), img.ctypes.data if len(img) else None, ctypes.byref(m),
                stream)


def load(fmt, path, stream=None):
    """binary_io<F>::load."""
    init()
    return _new(lib().ms_col_load, ctypes.byref(fmt), os.fsencode(path), stream)


def morph(fmt, col, stream=None):
    return _new(lib().ms_morph, ctypes.byref(fmt), col.h, stream)


def morph_blockwise(fmt, col, codes, stream=None):
    """A blockwise_individual_f column of col's values with the given format
    code per block (ms_format_code) instead of the format's selector."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    return _new(lib().ms_morph_blockwise, ctypes.byref(fmt), col.h, codes.ctypes.data if len(codes) else None,
                len(codes), stream)


def select(cmp, val, out_fmt, col, lanes=8, stream=None):
    return _new(lib().ms_select, cmp, val, ctypes.byref(out_fmt), col.h, lanes, stream)


def between(cmp_lower, val_lower, cmp_upper, val_upper, out_fmt, col, lanes=8, stream=None):
    return _new(lib().ms_between, cmp_lower, val_lower, cmp_upper, val_upper, ctypes.byref(out_fmt), col.h, lanes,
                stream)


def project(out_fmt, data, pos, lanes=8, stream=None):
    return _new(lib().ms_project, ctypes.byref(out_fmt), data.h, pos.h, lanes, stream)


def agg_sum(col, stream=None):
    return _new(lib().ms_agg_sum, col.h, stream)


def intersect_sorted(out_fmt, l, r, stream=None):
    return _new(lib().ms_intersect_sorted, ctypes.byref(out_fmt), l.h, r.h, stream)


def merge_sorted(out_fmt, l, r, stream=None):
    return _new(lib().ms_merge_sorted, ctypes.byref(out_fmt), l.h, r.h, stream)


def calc_binary(op, out_fmt, l, r, stream=None):
    return _new(lib().ms_calc_binary, op, ctypes.byref(out_fmt), l.h, r.h, stream)


def _new2(fn, *args):
    a, b = VP(), VP()
    check(fn(ctypes.byref(a), ctypes.byref(b), *args))
    st = next((x for x in args if isinstance(x, Stream)), None)
    return Column(a, st), Column(b, st)


def group(gr_fmt, ext_fmt, data, in_gr=None, stream=None):
    """group (unary without in_gr, binary with it): (group ids, extents)."""
    return _new2(lib().ms_group, ctypes.byref(gr_fmt), ctypes.byref(ext_fmt), in_gr.h if in_gr else None, data.h,
                 stream)


def join(l_fmt, r_fmt, l, r, stream=None):
    """nested_loop_join: (left positions, right positions) of equal pairs."""
    return _new2(lib().ms_join, ctypes.byref(l_fmt), ctypes.byref(r_fmt), l.h, r.h, stream)


def semi_join(out_fmt, l, r, stream=None):
    """left_semi_nto1_nested_loop_join: left positions with a match."""
    return _new(lib().ms_semi_join, ctypes.byref(out_fmt), l.h, r.h, stream)


class DeviceU64:
    """A device uint64 (for ms_select_sum results), via the HIP runtime."""

    _hip = None

    def __init__(self, n=1):
        if DeviceU64._hip is None:
            DeviceU64._hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        self.p = VP()
        self._owner = None
        rc = DeviceU64._hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(8 * n))
        if rc != 0:
            raise MsError(3, f"hipMalloc failed ({rc})")

    def at(self, k):
        """A view of element k (for ms_select_sum's result argument)."""
        v = object.__new__(DeviceU64)
        v.n, v.p, v._owner = 1, VP(self.p.value + 8 * k), self
        return v

    def zero(self):
        DeviceU64._hip.hipDeviceSynchronize()
        rc = DeviceU64._hip.hipMemset(self.p, 0, ctypes.c_size_t(8 * self.n))
        if rc != 0:
            raise MsError(3, f"hipMemset failed ({rc})")

    def get(self):
        out = np.zeros(self.n, np.uint64)
        DeviceU64._hip.hipDeviceSynchronize()
        rc = DeviceU64._hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.p, ctypes.c_size_t(8 * self.n), 2)
        if rc != 0:
            raise MsError(3, f"hipMemcpy failed ({rc})")
        return out

    def __del__(self):
        if self.p and DeviceU64._hip is not None and getattr(self, "_owner", None) is None:
            DeviceU64._hip.hipFree(self.p)


def select_sum(cmp, val, col, result=None, stream=None):
    """Fused select + agg_sum; returns the DeviceU64 holding the sum."""
    r = result or DeviceU64()
    check(lib().ms_select_sum(r.p, cmp, val, col.h, stream))
    return r


def select_sum_noalloc(cmp, val, col, result=None, stream=None):
    """ms_select_sum_noalloc: the fused select + agg_sum with no device
    allocation (its temporaries live in the library's per-device array)."""
    r = result or DeviceU64()
    check(lib().ms_select_sum_noalloc(r.p, cmp, val, col.h, stream))
    return r


class Plan:
    """A query plan recorded on `stream` and replayed as a hipGraph.

    with plan.record(stream): ...operators on `stream`...   (first: measure)
    with plan.record(stream): ...same operators...          (then: capture)
    plan.launch(stream)
    """

    def __init__(self, arena_bytes=0):
        init()
        self.h = VP()
        check(lib().ms_plan_create(ctypes.byref(self.h), arena_bytes))

    def record(self, stream):
        return _PlanRecording(self, stream)

    def launch(self, stream):
        check(lib().ms_plan_launch(self.h, stream))

    @property
    def arena_bytes(self):
        return lib().ms_plan_arena_bytes(self.h)

    def free(self):
        if self.h:
            lib().ms_plan_free(self.h)
            self.h = None

    def __del__(self):
        if self.h and _lib is not None:
            _lib.ms_plan_free(self.h)
            self.h = None


class _PlanRecording:
    """The context of Plan.record (a plain object: no reference cycle keeps
    the plan or the stream alive after the block)."""

    def __init__(self, plan, stream):
        self.plan, self.stream = plan, stream

    def __enter__(self):
        check(lib().ms_plan_begin(self.plan.h, self.stream))

    def __exit__(self, exc_type, *exc):
        rc = lib().ms_plan_end(self.plan.h, self.stream)
        if exc_type is None:        # an operator's own error takes precedence
            check(rc)
        return False


class Batch:
    """Fused select + sum over many columns in one persistent launch
    (ms_batch_*).  Keeps the enqueued columns alive."""

    def __init__(self):
        init()
        self.h = VP()
        check(lib().ms_batch_create(ctypes.byref(self.h)))
        self.cols = []

    def enqueue(self, col, cmp, val):
        e = U64()
        check(lib().ms_batch_enqueue(self.h, col.h, cmp, val, ctypes.byref(e)))
        self.cols.append(col)
        return e.value

    def run(self, results, stream=None):
        check(lib().ms_batch_run(self.h, results.p, stream))

    def view_times(self, results, reps=3, stream=None):
        """ns per view (entry, or block of a blockwise_individual_f entry):
        the wall-clock time of its work items in a timed run."""
        n = U64()
        check(lib().ms_batch_view_times(self.h, results.p, reps, None, 0, ctypes.byref(n), stream))
        out = np.zeros(max(n.value, 1), np.float64)
        check(lib().ms_batch_view_times(self.h, results.p, reps, out.ctypes.data, n.value, ctypes.byref(n), stream))
        return out[:n.value]

    def tune(self, results, reps=5, stream=None):
        c = BatchConfig()
        check(lib().ms_batch_tune(self.h, results.p, reps, ctypes.byref(c), stream))
        return c

    @property
    def config(self):
        c = BatchConfig()
        check(lib().ms_batch_get_config(self.h, ctypes.byref(c)))
        return c

    @config.setter
    def config(self, c):
        check(lib().ms_batch_set_config(self.h, ctypes.byref(c)))

    def timing(self, slots):
        """Record the persistent kernel's duration of the next runs (ring of
        `slots`; 0: off)."""
        check(lib().ms_batch_timing(self.h, slots))

    def kernel_times(self, n):
        ms = np.zeros(max(n, 1), np.float32)
        cnt = U32(n)
        check(lib().ms_batch_kernel_times(self.h, ms.ctypes.data, ctypes.byref(cnt)))
        return ms[:cnt.value].tolist()

    def stats(self):
        e, i, b = U64(), U64(), U64()
        check(lib().ms_batch_stats(self.h, ctypes.byref(e), ctypes.byref(i), ctypes.byref(b)))
        return e.value, i.value, b.value

    def __len__(self):
        return len(self.cols)

    def free(self):
        """Releases the batch (refused while a live plan has recorded it)."""
        if self.h:
            check(lib().ms_batch_free(self.h))
            self.h = None

    def __del__(self):
        if self.h and _lib is not None:
            _lib.ms_batch_free(self.h)
            self.h = None
