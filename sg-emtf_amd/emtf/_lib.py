"""ctypes binding of libemtf.so (include/emtf.h).

The HIP library is the only implementation: importing this module on a
machine without the built library raises, there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("EMTF_LIB") or os.path.join(os.path.dirname(_HERE), "libemtf.so")

EMTF_OK = 0
ERRORS = {-1: "EMTF_E_INVALID", -2: "EMTF_E_DEVICE", -3: "EMTF_E_NOMEM", -4: "EMTF_E_UNSUPPORTED"}

f32p = ctypes.POINTER(ctypes.c_float)


class TaskDesc(ctypes.Structure):
    _fields_ = [("func", ctypes.c_int32), ("dim", ctypes.c_int32), ("pop", ctypes.c_int32),
                ("xform", ctypes.c_int32), ("lo", ctypes.c_float), ("hi", ctypes.c_float),
                ("rot", f32p), ("shift", f32p), ("diag", f32p)]


class Params(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int32), ("F", ctypes.c_float), ("CR", ctypes.c_float),
                ("eta_c", ctypes.c_float), ("eta_m", ctypes.c_float), ("pm", ctypes.c_float),
                ("n_islands", ctypes.c_int32), ("exchange_interval", ctypes.c_int32),
                ("migrate_interval", ctypes.c_int32), ("migrate_count", ctypes.c_int32),
                ("transfer_interval", ctypes.c_int32), ("transfer_count", ctypes.c_int32),
                ("transfer_cr", ctypes.c_float), ("seed", ctypes.c_uint64), ("polish_interval", ctypes.c_int32),
                ("polish_steps", ctypes.c_int32), ("polish_topk", ctypes.c_int32), ("schedule", ctypes.c_int32), ("tune", ctypes.c_int32),
                ("formulas", ctypes.c_int32), ("admission", ctypes.c_int32)]


class EmtfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libemtf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run sg-emtf_amd/build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.emtf_create.argtypes = [ctypes.POINTER(TaskDesc), ctypes.c_int32, ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.emtf_run.argtypes = [vp, ctypes.c_int32, vp]
        L.emtf_admit.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp]
        L.emtf_sync.argtypes = [vp]
        L.emtf_generation.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.emtf_get_population.argtypes = [vp, ctypes.c_int32, f32p, f32p]
        L.emtf_set_population.argtypes = [vp, ctypes.c_int32, f32p, f32p]
        L.emtf_get_best.argtypes = [vp, f32p, f32p, ctypes.c_int32]
        L.emtf_get_elites.argtypes = [vp, ctypes.c_int32, f32p, f32p, ctypes.c_int32]
        L.emtf_trace_start.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp]
        L.emtf_trace.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), f32p, f32p]
        L.emtf_destroy.argtypes = [vp]
        L.emtf_last_error.restype = ctypes.c_char_p
        for n in ["emtf_create", "emtf_run", "emtf_admit", "emtf_sync", "emtf_generation", "emtf_get_population",
                  "emtf_set_population", "emtf_get_best", "emtf_get_elites", "emtf_trace_start", "emtf_trace",
                  "emtf_destroy"]:
            getattr(L, n).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != EMTF_OK:
        raise EmtfError(rc, lib().emtf_last_error().decode())
    return rc
