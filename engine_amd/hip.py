"""Minimal HIP runtime access through ctypes: events and synchronisation
for timing the C-ABI calls on the stream they run on."""
import ctypes

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with hipError {rc}")


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.e)), "hipEventCreate")

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.e, stream), "hipEventRecord")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _ok(hip().hipEventSynchronize(later.e), "hipEventSynchronize")
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, later.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if self.e and _hip is not None:
            _hip.hipEventDestroy(self.e)


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A created HIP stream (default priority).

    The object itself is what `handle` returns and what ctypes passes as the
    hipStream_t (`_as_parameter_`): the library's column wrappers keep the
    Stream they were created on alive, because a column is released in
    order on that stream, which therefore must outlive it."""

    def __init__(self):
        self.s = ctypes.c_void_p()
        self._columns = 0          # live columns created on this stream
        self._orphaned = False     # finalised while columns still used it
        _ok(hip().hipStreamCreate(ctypes.byref(self.s)), "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    @property
    def handle(self):
        return self

    @property
    def _as_parameter_(self):
        return self.s

    def _destroy(self):
        if self.s and _hip is not None:
            _hip.hipStreamDestroy(self.s)
        self.s = ctypes.c_void_p()

    def _column_released(self):
        self._columns -= 1
        if self._orphaned and self._columns <= 0:
            self._destroy()

    def __del__(self):
        # In a reference cycle the finalisers run in no particular order: a
        # column released later still frees its memory in order on this
        # stream, so the stream outlives the last such column.
        if self._columns > 0:
            self._orphaned = True
        else:
            self._destroy()


class DeviceArray:
    """A device buffer of n elements of a numpy dtype (hipMalloc)."""

    def __init__(self, n, dtype):
        import numpy as np
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(1, self.n * self.dtype.itemsize))),
            "hipMalloc")

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a, offset=0):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert offset + a.size <= self.n
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(ctypes.c_void_p(self.p.value + offset * self.dtype.itemsize),
                            ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")

    def numpy(self, count=None):
        import numpy as np
        count = self.n if count is None else count
        out = np.zeros(count, self.dtype)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.p, ctypes.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def ptr(self, offset=0):
        return ctypes.c_void_p(self.p.value + offset * self.dtype.itemsize)

    def __del__(self):
        if self.p and _hip is not None:
            _hip.hipFree(self.p)
            self.p = None
