"""HIP event timing on an arbitrary stream (the library's own stream).

torch.cuda.Event only sees torch's current stream; the solver kernels run on
the context's stream, so events are recorded there directly through
libamdhip64.
"""
from __future__ import annotations

import ctypes as C

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
    return _hip


class EventTimeline:
    """n+1 events recorded on `stream`; durations()[k] = ms between mark k and k+1."""

    def __init__(self, stream: int, n_marks: int):
        self.stream = C.c_void_p(stream)
        self.events = []
        for _ in range(n_marks):
            e = C.c_void_p()
            rc = hip().hipEventCreate(C.byref(e))
            if rc:
                raise RuntimeError(f"hipEventCreate failed ({rc})")
            self.events.append(e)
        self.k = 0

    def mark(self) -> None:
        rc = hip().hipEventRecord(self.events[self.k], self.stream)
        if rc:
            raise RuntimeError(f"hipEventRecord failed ({rc})")
        self.k += 1

    def durations(self) -> list[float]:
        hip().hipEventSynchronize(self.events[self.k - 1])
        out = []
        for a, b in zip(self.events[: self.k - 1], self.events[1: self.k]):
            ms = C.c_float()
            hip().hipEventElapsedTime(C.byref(ms), a, b)
            out.append(ms.value)
        return out

    def close(self) -> None:
        for e in self.events:
            hip().hipEventDestroy(e)
        self.events = []
