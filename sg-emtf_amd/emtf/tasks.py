"""Seeded task generators for the benchmark workloads.

The reference generates task parameters on the device (para_init,
proposed model/SG-EMTF.cu:181-195: M ~ U(0,1], b ~ U(0,1] * range, type
rand()%7).  The benchmark workloads instead fix, per task t:
  objective  [sphere, rastrigin, ackley, griewank, rosenbrock, schwefel,
              weierstrass, elliptic][t % 8]
  shift      o ~ U[-80, 80]^D
  rotation   Q from the QR factorisation of an N(0,1) D x D matrix with
             diag(R) > 0 (ROTATE workloads only)
  box        [-100, 100]^D
generated in float64 with numpy's PCG64 and stored as float32.
"""
import numpy as np

from .engine import DIAG, FUNC_IDS, ROTATE, SHIFT

CYCLE = ["sphere", "rastrigin", "ackley", "griewank", "rosenbrock", "schwefel", "weierstrass", "elliptic"]
# reference search boxes per objective id (rangeInit, SG-EMTF.cu:197-214)
REF_BOX = {0: (-100.0, 100.0), 1: (-50.0, 50.0), 2: (-50.0, 50.0), 3: (-50.0, 50.0),
           4: (-100.0, 100.0), 5: (-0.5, 0.5), 6: (-500.0, 500.0)}


def rotation(D, rng):
    A = rng.standard_normal((D, D))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))[None, :]
    return Q


def benchmark_tasks(n_tasks, dim, pop, rotate=True, seed=0, dims=None, pops=None, funcs=None):
    """Tasks cycling the eight objectives; dims/pops/funcs override per task."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tasks = []
    for t in range(n_tasks):
        D = int(dims[t]) if dims is not None else dim
        P = int(pops[t]) if pops is not None else pop
        f = int(funcs[t]) if funcs is not None else FUNC_IDS[CYCLE[t % 8]]
        shift = rng.uniform(-80.0, 80.0, D).astype(np.float32)
        task = dict(func=f, dim=D, pop=P, lo=-100.0, hi=100.0, shift=shift, xform=SHIFT)
        if rotate:
            task["rot"] = rotation(D, rng).astype(np.float32)
            task["xform"] = ROTATE
        tasks.append(task)
    return tasks


def uneven_tasks(n_tasks, seed=0, dmin=10, dmax=200, pmin=32, pmax=256):
    """Batched-uneven workload: D ~ U{dmin..dmax}, pop ~ U{pmin..pmax} rounded to 32."""
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    dims = rng.integers(dmin, dmax + 1, n_tasks)
    pops = np.maximum(32, (rng.integers(pmin, pmax + 1, n_tasks) + 16) // 32 * 32)
    return benchmark_tasks(n_tasks, 0, 0, rotate=True, seed=seed, dims=dims, pops=pops)


def reference_tasks(n_tasks, dim, pop, seed=0):
    """Reference-style tasks: z = M (x + b), M ~ U(0,1], b ~ U(0,1] * range, per-objective box."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tasks = []
    for t in range(n_tasks):
        f = int(rng.integers(0, 7))
        lo, hi = REF_BOX[f]
        M = (1.0 - rng.random(dim)).astype(np.float32)
        b = ((1.0 - rng.random(dim)) * (hi - lo)).astype(np.float32)
        tasks.append(dict(func=f, dim=dim, pop=pop, lo=lo, hi=hi, xform=DIAG, diag=M, shift=-b))
    return tasks
