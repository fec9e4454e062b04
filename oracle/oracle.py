"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py; the product never imports it.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

FUNC_NAMES = ["sphere", "rosenbrock", "ackley", "rastrigin", "griewank", "weierstrass", "schwefel", "elliptic"]
# the reference's own formulas (or_objectives.c): ids 8..13
REF_FUNC_IDS = {"ackley": 8, "rastrigin": 9, "griewank": 10, "weierstrass": 11, "schwefel": 12, "griewank_ga": 13}
XFORM_SHIFT, XFORM_ROTATE, XFORM_DIAG = 0, 1, 2
ALGO_DE_RAND_1_BIN, ALGO_DE_CURRENT_1_BIN, ALGO_GA_SBX_PM = 0, 1, 2


class OrTask(ctypes.Structure):
    _fields_ = [("func", ctypes.c_int32), ("dim", ctypes.c_int32), ("pop", ctypes.c_int32),
                ("xform", ctypes.c_int32), ("lo", ctypes.c_float), ("hi", ctypes.c_float),
                ("rot", ctypes.POINTER(ctypes.c_float)), ("shift", ctypes.POINTER(ctypes.c_float)),
                ("diag", ctypes.POINTER(ctypes.c_float))]


class OrParams(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int32), ("F", ctypes.c_float), ("CR", ctypes.c_float),
                ("eta_c", ctypes.c_float), ("eta_m", ctypes.c_float), ("pm", ctypes.c_float),
                ("n_islands", ctypes.c_int32), ("exchange_interval", ctypes.c_int32),
                ("migrate_interval", ctypes.c_int32), ("migrate_count", ctypes.c_int32),
                ("transfer_interval", ctypes.c_int32), ("transfer_count", ctypes.c_int32),
                ("transfer_cr", ctypes.c_float), ("seed", ctypes.c_uint64), ("polish_interval", ctypes.c_int32),
                ("polish_steps", ctypes.c_int32), ("polish_topk", ctypes.c_int32)]


class CvrpInstance(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("capacity", ctypes.c_int32), ("pop", ctypes.c_int32),
                ("xy", ctypes.POINTER(ctypes.c_int32)), ("demand", ctypes.POINTER(ctypes.c_int32))]


class CvrpParams(ctypes.Structure):
    _fields_ = [("p_swap", ctypes.c_float), ("p_2opt", ctypes.c_float), ("n_islands", ctypes.c_int32),
                ("migrate_interval", ctypes.c_int32), ("migrate_count", ctypes.c_int32), ("seed", ctypes.c_uint64)]


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)


_lib = None


def build_mutant(k, out):
    """Builds the oracle with planted mistake OR_MUTANT=k into `out` (mutation tests only)."""
    subprocess.run(["make", "-s", "-C", _HERE, "mutant", f"K={int(k)}", f"OUT={out}"], check=True)
    return out


@contextlib.contextmanager
def swapped(path):
    """Routes every call made inside the block through another build of the oracle (a
    mutant): objects created inside the block keep using it."""
    global _lib
    old = _lib
    _lib = _bind(path)
    try:
        yield _lib
    finally:
        _lib = old


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = _bind(_LIB)
    return _lib


def _bind(path):
    L = ctypes.CDLL(path)
    f32p = ctypes.POINTER(ctypes.c_float)
    f64p = ctypes.POINTER(ctypes.c_double)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.or_philox.argtypes = [u32p, u32p, u32p]
    L.or_draw_block.argtypes = [ctypes.c_uint64] + [ctypes.c_uint32] * 6 + [u32p]
    for n in ["or_sinpi_sq", "or_sin2pi", "or_expm1", "or_exp2", "or_log2", "or_weier_term", "or_cospi", "or_exp",
              "or_weier_ref_term"]:
        getattr(L, n).argtypes = [ctypes.c_float]
        getattr(L, n).restype = ctypes.c_float
    L.or_weier_ref_w.argtypes = []
    L.or_weier_ref_w.restype = ctypes.c_float
    L.or_pow.argtypes = [ctypes.c_float, ctypes.c_float]
    L.or_pow.restype = ctypes.c_float
    L.or_func_table.argtypes = [ctypes.c_int32, ctypes.c_int32, f32p]
    L.or_eval32.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, f32p, f32p]
    L.or_eval64.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, f64p, f64p]
    L.or_grad64.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, f64p, f64p]
    L.or_grad32.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, f32p, f32p]
    L.or_create.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, ctypes.POINTER(OrParams)]
    L.or_create.restype = ctypes.c_void_p
    L.or_create_deferred.argtypes = [ctypes.POINTER(OrTask), ctypes.c_int32, ctypes.POINTER(OrParams)]
    L.or_create_deferred.restype = ctypes.c_void_p
    L.or_admit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.or_admit.restype = ctypes.c_int32
    L.or_transfer_donor.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.or_transfer_donor.restype = ctypes.c_int32
    L.or_destroy.argtypes = [ctypes.c_void_p]
    L.or_run.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.or_generation.argtypes = [ctypes.c_void_p]
    L.or_generation.restype = ctypes.c_int32
    L.or_get_population.argtypes = [ctypes.c_void_p, ctypes.c_int32, f32p, f32p]
    L.or_set_population.argtypes = [ctypes.c_void_p, ctypes.c_int32, f32p, f32p]
    L.or_get_best.argtypes = [ctypes.c_void_p, f32p, f32p, ctypes.c_int32]
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.or_cvrp_dist.argtypes = [ctypes.c_int32] * 4
    L.or_cvrp_dist.restype = ctypes.c_int32
    L.or_cvrp_split.argtypes = [u8p, ctypes.c_int, u8p, u8p, ctypes.c_int]
    L.or_cvrp_split.restype = ctypes.c_int32
    L.or_cvrp_child.argtypes = [u8p, u8p] + [ctypes.c_int] * 6 + [u8p]
    L.or_cvrp_create.argtypes = [ctypes.POINTER(CvrpInstance), ctypes.c_int32, ctypes.POINTER(CvrpParams)]
    L.or_cvrp_create.restype = ctypes.c_void_p
    L.or_cvrp_run.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.or_cvrp_get.argtypes = [ctypes.c_void_p, ctypes.c_int32, u8p, i32p]
    L.or_cvrp_destroy.argtypes = [ctypes.c_void_p]
    L.or_phase_exchange.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    L.or_phase_migrate.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.or_phase_transfer.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.or_cvrp_set.argtypes = [ctypes.c_void_p, ctypes.c_int32, u8p, i32p]
    L.or_cvrp_phase_migrate.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    return L


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox(c, k, o)
    return list(o)


def draw_block(seed, task, gen, purpose, index, sub, block):
    o = (ctypes.c_uint32 * 4)()
    lib().or_draw_block(seed, task, gen, purpose, index, sub, block, o)
    return list(o)


def scalar(name, *args):
    return getattr(lib(), "or_" + name)(*args)


def func_table(func, dim):
    t = np.zeros(dim, np.float32)
    lib().or_func_table(func, dim, _fp(t))
    return t


class _TaskHolder:
    """Keeps the numpy arrays of one task alive while C holds pointers."""

    def __init__(self, task):
        self.func = int(task["func"])
        self.dim = int(task["dim"])
        self.pop = int(task.get("pop", 32))
        self.xform = int(task.get("xform", XFORM_SHIFT))
        self.lo = float(task.get("lo", -100.0))
        self.hi = float(task.get("hi", 100.0))
        D = self.dim
        self.shift = np.ascontiguousarray(task.get("shift", np.zeros(D)), dtype=np.float32)
        self.rot = np.ascontiguousarray(task["rot"], dtype=np.float32) if task.get("rot") is not None else None
        self.diag = np.ascontiguousarray(task["diag"], dtype=np.float32) if task.get("diag") is not None else None
        self.c = OrTask(self.func, D, self.pop, self.xform, self.lo, self.hi,
                        _fp(self.rot) if self.rot is not None else None, _fp(self.shift),
                        _fp(self.diag) if self.diag is not None else None)


def eval32(task, x):
    h = _TaskHolder(task)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, h.dim)
    f = np.zeros(x.shape[0], np.float32)
    lib().or_eval32(ctypes.byref(h.c), x.shape[0], _fp(x), _fp(f))
    return f


def eval64(task, x):
    h = _TaskHolder(task)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, h.dim)
    f = np.zeros(x.shape[0], np.float64)
    lib().or_eval64(ctypes.byref(h.c), x.shape[0], _dp(x), _dp(f))
    return f


def grad32(task, x):
    h = _TaskHolder(task)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, h.dim)
    g = np.zeros_like(x)
    lib().or_grad32(ctypes.byref(h.c), x.shape[0], _fp(x), _fp(g))
    return g


def grad64(task, x):
    h = _TaskHolder(task)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, h.dim)
    g = np.zeros_like(x)
    lib().or_grad64(ctypes.byref(h.c), x.shape[0], _dp(x), _dp(g))
    return g


def make_params(**kw):
    d = dict(algo=ALGO_DE_RAND_1_BIN, F=0.5, CR=0.9, eta_c=2.0, eta_m=5.0, pm=1.0, n_islands=1,
             exchange_interval=0, migrate_interval=0, migrate_count=0, transfer_interval=0,
             transfer_count=0, transfer_cr=0.0, seed=0, polish_interval=0, polish_steps=0, polish_topk=0)
    d.update(kw)
    return OrParams(**d)


class Engine:
    """The oracle's multitasking engine over a list of task dicts."""

    def __init__(self, tasks, deferred=False, **params):
        """deferred: online admission, no task runs before admit() starts it."""
        self._holders = [_TaskHolder(t) for t in tasks]
        arr = (OrTask * len(tasks))(*[h.c for h in self._holders])
        self._p = make_params(**params)
        self._L = lib()
        create = self._L.or_create_deferred if deferred else self._L.or_create
        self._h = create(arr, len(tasks), ctypes.byref(self._p))
        self.n_tasks = len(tasks)

    def admit(self, first, count, generations):
        """Start tasks first .. first + count - 1 now for `generations` generations."""
        if self._L.or_admit(self._h, int(first), int(count), int(generations)) != 0:
            raise ValueError("bad admission (range, repeated task, or off a transfer boundary)")

    def transfer_donor(self, t, clock):
        """The donor task t draws at the transfer boundary `clock` (-1: none)."""
        return self._L.or_transfer_donor(self._h, int(t), int(clock))

    def run(self, gens, threads=0):
        self._L.or_run(self._h, int(gens), int(threads))

    @property
    def generation(self):
        return self._L.or_generation(self._h)

    def population(self, t):
        h = self._holders[t]
        x = np.zeros((h.pop, h.dim), np.float32)
        f = np.zeros(h.pop, np.float32)
        self._L.or_get_population(self._h, t, _fp(x), _fp(f))
        return x, f

    def set_population(self, t, x, f):
        x = np.ascontiguousarray(x, np.float32)
        f = np.ascontiguousarray(f, np.float32)
        self._L.or_set_population(self._h, t, _fp(x), _fp(f))

    def best(self):
        dmax = max(h.dim for h in self._holders)
        f = np.zeros(self.n_tasks, np.float32)
        x = np.zeros((self.n_tasks, dmax), np.float32)
        self._L.or_get_best(self._h, _fp(f), _fp(x), dmax)
        return f, x

    # single phases on the current state (tests that pin them one at a time)
    def phase_exchange(self, t, gen):
        self._L.or_phase_exchange(self._h, int(t), int(gen))

    def phase_migrate(self, t):
        self._L.or_phase_migrate(self._h, int(t))

    def phase_transfer(self, gen):
        self._L.or_phase_transfer(self._h, int(gen))

    def close(self):
        if self._h:
            self._L.or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------- CVRP
def _u8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def cvrp_dist_matrix(xy):
    n1 = xy.shape[0]
    d = np.zeros((n1, n1), np.uint8)
    for i in range(n1):
        for j in range(n1):
            d[i, j] = lib().or_cvrp_dist(int(xy[i, 0]), int(xy[i, 1]), int(xy[j, 0]), int(xy[j, 1]))
    return d


def cvrp_split(tour, dist, demand, cap):
    t = np.ascontiguousarray(tour, np.uint8)
    return lib().or_cvrp_split(_u8(t), len(t), _u8(np.ascontiguousarray(dist, np.uint8)),
                               _u8(np.ascontiguousarray(demand, np.uint8)), int(cap))


def cvrp_child(p1, p2, a, b, mut, c, d):
    p1 = np.ascontiguousarray(p1, np.uint8)
    p2 = np.ascontiguousarray(p2, np.uint8)
    out = np.zeros_like(p1)
    lib().or_cvrp_child(_u8(p1), _u8(p2), len(p1), a, b, mut, c, d, _u8(out))
    return out


class CvrpEngine:
    """The oracle's CVRP GA over a list of instance dicts (n, capacity, pop, xy, demand)."""

    def __init__(self, instances, p_swap=0.3, p_2opt=0.3, n_islands=1, migrate_interval=0, migrate_count=0, seed=0):
        self._keep = []
        arr = []
        for ins in instances:
            xy = np.ascontiguousarray(ins["xy"], np.int32)
            dem = np.ascontiguousarray(ins["demand"], np.int32)
            self._keep += [xy, dem]
            arr.append(CvrpInstance(int(ins["n"]), int(ins["capacity"]), int(ins["pop"]), _i32(xy), _i32(dem)))
        carr = (CvrpInstance * len(arr))(*arr)
        self._p = CvrpParams(p_swap, p_2opt, n_islands, migrate_interval, migrate_count, seed)
        self._L = lib()
        self._h = self._L.or_cvrp_create(carr, len(arr), ctypes.byref(self._p))
        self.instances = instances

    def transfer_donor(self, t, clock):
        """The donor task t draws at the transfer boundary `clock` (-1: none)."""
        return self._L.or_transfer_donor(self._h, int(t), int(clock))

    def run(self, gens, threads=0):
        self._L.or_cvrp_run(self._h, int(gens), int(threads))

    def population(self, i):
        ins = self.instances[i]
        t = np.zeros((ins["pop"], ins["n"]), np.uint8)
        c = np.zeros(ins["pop"], np.int32)
        self._L.or_cvrp_get(self._h, i, _u8(t), _i32(c))
        return t, c

    def set_population(self, i, tours, costs):
        t = np.ascontiguousarray(tours, np.uint8)
        c = np.ascontiguousarray(costs, np.int32)
        self._L.or_cvrp_set(self._h, i, _u8(t), _i32(c))

    def phase_migrate(self, i):
        self._L.or_cvrp_phase_migrate(self._h, int(i))

    def close(self):
        if self._h:
            self._L.or_cvrp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
