"""Minimal HIP runtime access (events, device sync) through ctypes."""
import ctypes

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = ctypes.CDLL("libamdhip64.so")
        h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        h.hipEventDestroy.argtypes = [ctypes.c_void_p]
        h.hipEventQuery.argtypes = [ctypes.c_void_p]
        h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        h.hipGetErrorString.restype = ctypes.c_char_p
        _hip = h
    return _hip


def _ok(rc):
    if rc != 0:
        raise RuntimeError("HIP error: " + hip().hipGetErrorString(rc).decode())


def device_synchronize():
    _ok(hip().hipDeviceSynchronize())


class Event:
    def __init__(self):
        self.e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(self.e)))

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.e, ctypes.c_void_p(stream or 0)))

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.e))

    def query(self):
        """True once the work recorded before the event has finished."""
        rc = hip().hipEventQuery(self.e)
        if rc == 600:  # hipErrorNotReady
            return False
        _ok(rc)
        return True

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, end.e))
        return ms.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


HIP_STREAM_NON_BLOCKING = 0x1


class Stream:
    """A non-blocking stream at the default priority (no implicit ordering with the null stream)."""

    def __init__(self):
        self.s = ctypes.c_void_p()
        _ok(hip().hipStreamCreateWithFlags(ctypes.byref(self.s), HIP_STREAM_NON_BLOCKING))

    @property
    def handle(self):
        return self.s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.s))

    def __del__(self):
        try:
            hip().hipStreamDestroy(self.s)
        except Exception:
            pass
