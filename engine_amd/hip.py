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
    """A created HIP stream (default priority)."""

    def __init__(self):
        self.s = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.s)), "hipStreamCreate")

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.s), "hipStreamSynchronize")

    @property
    def handle(self):
        return self.s

    def __del__(self):
        if self.s and _hip is not None:
            _hip.hipStreamDestroy(self.s)
