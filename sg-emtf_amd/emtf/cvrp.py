"""CVRP genetic algorithm over the C ABI (emtf_cvrp_* in include/emtf.h).

Replaces the reference's CVRP variant (proposed model/SG-EMTF_cvrp.cu); see
sg-emtf_amd/csrc/cvrp.hip for the algorithm.
"""
import ctypes

import numpy as np

from . import _lib

i32p = ctypes.POINTER(ctypes.c_int32)
u8p = ctypes.POINTER(ctypes.c_uint8)


class CvrpInstance(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("capacity", ctypes.c_int32), ("pop", ctypes.c_int32),
                ("xy", i32p), ("demand", i32p)]


class CvrpParams(ctypes.Structure):
    _fields_ = [("p_swap", ctypes.c_float), ("p_2opt", ctypes.c_float), ("n_islands", ctypes.c_int32),
                ("migrate_interval", ctypes.c_int32), ("migrate_count", ctypes.c_int32), ("seed", ctypes.c_uint64)]


def _bind():
    L = _lib.lib()
    if not getattr(L, "_cvrp_bound", False):
        vp = ctypes.c_void_p
        L.emtf_cvrp_create.argtypes = [ctypes.POINTER(CvrpInstance), ctypes.c_int32, ctypes.POINTER(CvrpParams),
                                       ctypes.POINTER(vp)]
        L.emtf_cvrp_run.argtypes = [vp, ctypes.c_int32, vp]
        L.emtf_cvrp_sync.argtypes = [vp]
        L.emtf_cvrp_get_population.argtypes = [vp, ctypes.c_int32, u8p, i32p]
        L.emtf_cvrp_get_best.argtypes = [vp, i32p, u8p, ctypes.c_int32]
        L.emtf_cvrp_destroy.argtypes = [vp]
        L._cvrp_bound = True
    return L


class CvrpEngine:
    """GPU CVRP GA over instance dicts: n, capacity, pop, xy[(n+1), 2], demand[n+1]."""

    def __init__(self, instances, p_swap=0.3, p_2opt=0.3, n_islands=1, migrate_interval=0, migrate_count=0, seed=0):
        L = _bind()
        self.instances = instances
        self._keep = []
        arr = []
        for ins in instances:
            xy = np.ascontiguousarray(ins["xy"], np.int32)
            dem = np.ascontiguousarray(ins["demand"], np.int32)
            self._keep += [xy, dem]
            arr.append(CvrpInstance(int(ins["n"]), int(ins["capacity"]), int(ins["pop"]),
                                    xy.ctypes.data_as(i32p), dem.ctypes.data_as(i32p)))
        carr = (CvrpInstance * len(arr))(*arr)
        prm = CvrpParams(p_swap, p_2opt, n_islands, migrate_interval, migrate_count, seed)
        h = ctypes.c_void_p()
        _lib.check(L.emtf_cvrp_create(carr, len(arr), ctypes.byref(prm), ctypes.byref(h)))
        self._h = h

    def run(self, generations, stream=None):
        _lib.check(_bind().emtf_cvrp_run(self._h, int(generations), ctypes.c_void_p(stream or 0)))

    def sync(self):
        _lib.check(_bind().emtf_cvrp_sync(self._h))

    def population(self, i):
        ins = self.instances[i]
        t = np.zeros((ins["pop"], ins["n"]), np.uint8)
        c = np.zeros(ins["pop"], np.int32)
        _lib.check(_bind().emtf_cvrp_get_population(self._h, i, t.ctypes.data_as(u8p), c.ctypes.data_as(i32p)))
        return t, c

    def best(self):
        nmax = max(int(i["n"]) for i in self.instances)
        c = np.zeros(len(self.instances), np.int32)
        t = np.zeros((len(self.instances), nmax), np.uint8)
        _lib.check(_bind().emtf_cvrp_get_best(self._h, c.ctypes.data_as(i32p), t.ctypes.data_as(u8p), nmax))
        return c, t

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().emtf_cvrp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_instances(n_inst, seed=0, nmin=50, nmax=200, pop=128, capacity=50):
    """Benchmark CVRP instances: n ~ U{nmin..nmax} customers at integer U[0,100]^2,
    depot (50, 50), demand U{1..10}, the given capacity."""
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    out = []
    for _ in range(n_inst):
        n = int(rng.integers(nmin, nmax + 1))
        xy = np.vstack([[50, 50], rng.integers(0, 101, (n, 2))]).astype(np.int32)
        dem = np.concatenate([[0], rng.integers(1, 11, n)]).astype(np.int32)
        out.append(dict(n=n, capacity=capacity, pop=pop, xy=xy, demand=dem))
    return out
