"""Host-side mirror of the reference's per-task knobs over the C ABI.

`Engine(tasks, **params)` corresponds to the reference's initialization() +
iter() pair (proposed model/SG-EMTF.cu:293-380, :1646-1745): it owns the
device populations of a batch of tasks and advances them by generations.
Task dicts carry the reference's per-task state (objective id, dimension,
population, box, transform); params carry its #define knobs.
"""
import ctypes

import numpy as np

from . import _lib

FUNC_IDS = {"sphere": 0, "rosenbrock": 1, "ackley": 2, "rastrigin": 3, "griewank": 4,
            "weierstrass": 5, "schwefel": 6, "elliptic": 7}
SHIFT, ROTATE, DIAG = 0, 1, 2
DE_RAND_1_BIN, DE_CURRENT_1_BIN, GA_SBX_PM = 0, 1, 2
SCHEDULE_AUTO, SCHEDULE_EPOCHS, SCHEDULE_PERSISTENT = 0, 1, 2
# objective formulas: this engine's binary32 forms, or the reference's own (its DE file's
# evaluation(), or the GA file's copy whose Griewank has no product)
FORMULAS_TEXTBOOK, FORMULAS_REFERENCE_DE, FORMULAS_REFERENCE_GA = 0, 1, 2
# when tasks start: at create, or online (each batch at its emtf_admit)
ADMIT_AT_CREATE, ADMIT_ONLINE = 0, 1

DEFAULT_PARAMS = dict(algo=DE_RAND_1_BIN, F=0.5, CR=0.9, eta_c=2.0, eta_m=5.0, pm=1.0, n_islands=1,
                      exchange_interval=0, migrate_interval=0, migrate_count=0, transfer_interval=0,
                      transfer_count=0, transfer_cr=0.0, seed=0, polish_interval=0, polish_steps=0, polish_topk=0,
                      schedule=0, tune=0, formulas=0, admission=0)


def _fp(a):
    return a.ctypes.data_as(_lib.f32p) if a is not None else None


class Engine:
    def __init__(self, tasks, **params):
        p = dict(DEFAULT_PARAMS)
        unknown = set(params) - set(p)
        if unknown:
            raise TypeError(f"unknown engine parameters {sorted(unknown)}")
        p.update(params)
        self.params = p
        self.tasks = tasks
        self._keep = []
        descs = []
        for t in tasks:
            D = int(t["dim"])
            shift = np.ascontiguousarray(t.get("shift", np.zeros(D)), dtype=np.float32)
            rot = t.get("rot")
            rot = np.ascontiguousarray(rot, dtype=np.float32) if rot is not None else None
            diag = t.get("diag")
            diag = np.ascontiguousarray(diag, dtype=np.float32) if diag is not None else None
            self._keep += [shift, rot, diag]
            descs.append(_lib.TaskDesc(int(t["func"]), D, int(t["pop"]), int(t.get("xform", SHIFT)),
                                       float(t.get("lo", -100.0)), float(t.get("hi", 100.0)),
                                       _fp(rot), _fp(shift), _fp(diag)))
        arr = (_lib.TaskDesc * len(descs))(*descs)
        prm = _lib.Params(**p)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().emtf_create(arr, len(descs), ctypes.byref(prm), ctypes.byref(h)))
        self._h = h
        self.n_tasks = len(tasks)
        self.max_dim = max(int(t["dim"]) for t in tasks)

    def run(self, generations, stream=None):
        """Queue `generations` generations (asynchronous on `stream`, an int handle or None)."""
        _lib.check(_lib.lib().emtf_run(self._h, int(generations), ctypes.c_void_p(stream or 0)))

    def admit(self, first, count, generations, stream=None):
        """Online admission: start tasks first .. first + count - 1 at the current clock for
        `generations` generations (asynchronous on `stream`)."""
        _lib.check(_lib.lib().emtf_admit(self._h, int(first), int(count), int(generations),
                                         ctypes.c_void_p(stream or 0)))

    def sync(self):
        _lib.check(_lib.lib().emtf_sync(self._h))

    @property
    def generation(self):
        g = ctypes.c_int32()
        _lib.check(_lib.lib().emtf_generation(self._h, ctypes.byref(g)))
        return g.value

    def population(self, t):
        D, P = int(self.tasks[t]["dim"]), int(self.tasks[t]["pop"])
        x = np.zeros((P, D), np.float32)
        f = np.zeros(P, np.float32)
        _lib.check(_lib.lib().emtf_get_population(self._h, t, _fp(x), _fp(f)))
        return x, f

    def set_population(self, t, x, f):
        x = np.ascontiguousarray(x, np.float32)
        f = np.ascontiguousarray(f, np.float32)
        _lib.check(_lib.lib().emtf_set_population(self._h, t, _fp(x), _fp(f)))

    def best(self):
        f = np.zeros(self.n_tasks, np.float32)
        x = np.zeros((self.n_tasks, self.max_dim), np.float32)
        _lib.check(_lib.lib().emtf_get_best(self._h, _fp(f), _fp(x), self.max_dim))
        return f, x

    def elites(self, k):
        """The k best individuals of every task: fitness [T, k] and x [T, k, max_dim] (zero past dim)."""
        f = np.zeros((self.n_tasks, k), np.float32)
        x = np.zeros((self.n_tasks, k, self.max_dim), np.float32)
        _lib.check(_lib.lib().emtf_get_elites(self._h, int(k), _fp(f), _fp(x), self.max_dim))
        return f, x

    def trace_start(self, interval, max_points, stream=None):
        """Record every task's best on the device now and every `interval` generations
        (at most max_points points); interval 0 stops recording."""
        _lib.check(_lib.lib().emtf_trace_start(self._h, int(interval), int(max_points), ctypes.c_void_p(stream or 0)))

    def trace(self):
        """The recorded trace as [n_tasks, points + 1] float32 rows: the best at each point,
        then the worst fitness at the first point (ConvergenceRecorder.rows' layout)."""
        n = ctypes.c_int32()
        _lib.check(_lib.lib().emtf_trace(self._h, ctypes.byref(n), None, None))
        best = np.zeros((max(n.value, 1), self.n_tasks), np.float32)
        worst = np.zeros(self.n_tasks, np.float32)
        _lib.check(_lib.lib().emtf_trace(self._h, ctypes.byref(n), _fp(best), _fp(worst)))
        return np.concatenate([best[: n.value].T, worst[:, None]], 1)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().emtf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
