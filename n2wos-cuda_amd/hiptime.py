"""HIP events through the runtime library (timing kernels on the stream they run on)."""
from __future__ import annotations

import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so")
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip = h
    return _hip


class EventPair:
    """Two hipEvent_t handles; elapsed() waits for the second and returns ms."""

    def __init__(self):
        self.begin = C.c_void_p()
        self.end = C.c_void_p()
        assert hip().hipEventCreate(C.byref(self.begin)) == 0
        assert hip().hipEventCreate(C.byref(self.end)) == 0

    @property
    def handles(self):
        return self.begin.value, self.end.value

    def elapsed(self) -> float:
        assert hip().hipEventSynchronize(self.end) == 0
        ms = C.c_float()
        assert hip().hipEventElapsedTime(C.byref(ms), self.begin, self.end) == 0
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.begin)
            hip().hipEventDestroy(self.end)
        except Exception:
            pass
