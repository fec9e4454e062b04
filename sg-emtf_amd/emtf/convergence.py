"""Convergence records, ANO statistics and the FP64 error of best points.

Record format (the reference's writer, proposed model/SG-EMTF_DE.cu:1793-1811):
one line per task holding the task's best fitness at generation 0 and after
every `interval` generations, followed by the worst fitness of the initial
population; each value is printed like a C++ ostream float (six significant
digits) followed by one space.  Runs are appended to the same file, each closed
by a blank line.

ANO, the average normalised objective of the reference's evaluation program
(data process/源.cpp:122-205): for task i, min_i is the smallest final best over
the runs and max_i the largest initial worst; point j of the curve is the mean
over tasks and runs of (v_kij - min_i) / (max_i - min_i), a task with
max_i == min_i contributing 0.  The sums run in binary32 in the order tasks,
points, runs, as the reference's program does, so its printed curves are
reproduced digit for digit.

error64 re-evaluates points in float64 with cancellation-free forms of the
textbook objectives (z = x - o, Q (x - o) or diag * (x - o)); every objective
has its minimum 0 at z = 0 (Rosenbrock: z = 1) except Schwefel, whose
constrained optimum lies outside the box for shifted / rotated tasks and which
therefore has no known minimum (NaN error).
"""
import math

import numpy as np

from .engine import DIAG, ROTATE

SCHWEFEL = 6


def _z64(task, x):
    x = np.asarray(x, np.float64).reshape(-1, int(task["dim"]))
    o = np.asarray(task.get("shift", np.zeros(x.shape[1])), np.float64)
    w = x - o[None, :]
    xf = int(task.get("xform", 0))
    if xf == ROTATE:
        return w @ np.asarray(task["rot"], np.float64).T
    if xf == DIAG:
        return w * np.asarray(task["diag"], np.float64)[None, :]
    return w


def f64_of_z(func, z):
    """f(z) in float64 for rows z[n, D] of one objective."""
    z = np.asarray(z, np.float64)
    n, D = z.shape
    j = np.arange(D, dtype=np.float64)
    f = int(func)
    if f == 0:
        return (z * z).sum(1)
    if f == 1:
        a = z[:, :-1] ** 2 - z[:, 1:]
        return (100.0 * a * a + (z[:, :-1] - 1.0) ** 2).sum(1)
    if f == 2:
        r = np.sqrt((z * z).sum(1) / D)
        s = (np.sin(np.pi * np.fmod(z, 1.0)) ** 2).sum(1)
        return -20.0 * np.expm1(-0.2 * r) - math.e * np.expm1(-2.0 * s / D)
    if f == 3:
        return (z * z + 20.0 * np.sin(np.pi * np.fmod(z, 1.0)) ** 2).sum(1)
    if f == 4:
        th = z / np.sqrt(j + 1.0)
        c = 1.0 - 2.0 * np.sin(0.5 * th) ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            one_minus_prod = np.where((c > 0).all(1), -np.expm1(np.log1p(c - 1.0).sum(1)), 1.0 - np.prod(c, 1))
        return (z * z).sum(1) / 4000.0 + one_minus_prod
    if f == 5:
        acc = np.zeros(n)
        for k in range(21):
            t = np.fmod(z * 3.0 ** k, 1.0)
            acc += 2.0 ** (1 - k) * (np.sin(np.pi * t) ** 2).sum(1)
        return acc
    if f == 6:
        return 418.9828872724338 * D - (z * np.sin(np.sqrt(np.abs(z)))).sum(1)
    if f == 7:
        c = 1e6 ** (j / (D - 1)) if D > 1 else np.ones(1)
        return (c[None, :] * z * z).sum(1)
    raise ValueError(f"unknown objective {f}")


def objective64(task, x):
    """f(z) in float64 for points x[n, dim] of one task."""
    return f64_of_z(task["func"], _z64(task, x))


def error64(tasks, x_best):
    """f64(best) - f* per task (NaN for Schwefel); x_best rows may be wider than dim.

    Tasks sharing objective and dimension are transformed and evaluated together."""
    err = np.full(len(tasks), np.nan)
    groups = {}
    for t, task in enumerate(tasks):
        if int(task["func"]) != SCHWEFEL:
            groups.setdefault((int(task["func"]), int(task["dim"]), int(task.get("xform", 0))), []).append(t)
    for (func, D, xf), ids in groups.items():
        x = np.stack([np.asarray(x_best[t][:D], np.float64) for t in ids])
        w = x - np.stack([np.asarray(tasks[t].get("shift", np.zeros(D)), np.float64) for t in ids])
        if xf == ROTATE:
            Q = np.stack([np.asarray(tasks[t]["rot"], np.float64) for t in ids])
            z = np.einsum("tij,tj->ti", Q, w)
        elif xf == DIAG:
            z = w * np.stack([np.asarray(tasks[t]["diag"], np.float64) for t in ids])
        else:
            z = w
        err[ids] = f64_of_z(func, z)
    return err


class ConvergenceRecorder:
    """Best fitness of every task every `interval` generations of an engine run."""

    def __init__(self, engine, interval):
        self.engine = engine
        self.interval = int(interval)
        worst = np.empty(engine.n_tasks, np.float32)
        for t in range(engine.n_tasks):
            _, f = engine.population(t)
            worst[t] = np.nanmax(np.where(np.isnan(f), -np.inf, f))
        self.initial_worst = worst
        self.bests = [engine.best()[0]]

    def run(self, generations):
        for _ in range(int(generations) // self.interval):
            self.engine.run(self.interval)
            self.bests.append(self.engine.best()[0])

    def rows(self):
        """[n_tasks, points + 1] float32: the best at each point, then the initial worst."""
        return np.concatenate([np.stack(self.bests, 1), self.initial_worst[:, None]], 1).astype(np.float32)


def write_para(path, tasks, rotations_path=None):
    """Task parameters in the reference's para.txt format (SG-EMTF.cu:244-264): one line per
    task of `M_j b_j` pairs per dimension, a blank line, then the objective ids.  The
    reference's transform is z = M (x + b) with diagonal M, so b = -o and M = the task's
    diagonal (1 for shift-only tasks).  Rotations do not fit the format: rotated tasks get
    M = 1 here and their Q matrices are written to `rotations_path` (.npz, key "task<t>"),
    which is then required."""
    rots = {}
    with open(path, "w") as fh:
        for t, task in enumerate(tasks):
            D = int(task["dim"])
            o = np.asarray(task.get("shift", np.zeros(D)), np.float32)
            xf = int(task.get("xform", 0))
            m = np.asarray(task["diag"], np.float32) if xf == DIAG else np.ones(D, np.float32)
            if xf == ROTATE:
                rots[f"task{t}"] = np.asarray(task["rot"], np.float32)
            fh.write("".join(f"{_fmt(m[j])} {_fmt(-o[j])} " for j in range(D)) + "\n")
        fh.write("\n")
        fh.write("".join(f"{int(task['func'])} " for task in tasks) + "\n")
    if rots:
        if rotations_path is None:
            raise ValueError("rotated tasks need rotations_path for their Q matrices")
        np.savez(rotations_path, **rots)


def read_para(path):
    """(M [T][D], b [T][D], objective ids [T]) from a para.txt file."""
    with open(path) as fh:
        blocks = fh.read().split("\n\n")
    rows = [[float(v) for v in line.split()] for line in blocks[0].strip().split("\n")]
    M = [np.asarray(r[0::2], np.float32) for r in rows]
    b = [np.asarray(r[1::2], np.float32) for r in rows]
    ids = [int(v) for v in blocks[1].split()]
    return M, b, ids


def _fmt(v):
    return "%g" % float(np.float32(v))


def write_record(path, rows, append=True):
    """Append one run (rows from ConvergenceRecorder.rows) in the reference's record format."""
    with open(path, "a" if append else "w") as fh:
        for r in np.asarray(rows, np.float32):
            fh.write("".join(_fmt(v) + " " for v in r) + "\n")
        fh.write("\n")


def read_record(path):
    """[runs, tasks, points + 1] float32 from a record file."""
    runs, cur = [], []
    with open(path) as fh:
        for line in fh:
            vals = line.split()
            if not vals:
                if cur:
                    runs.append(cur)
                cur = []
                continue
            cur.append([np.float32(float(v)) for v in vals])
    if cur:
        runs.append(cur)
    return np.asarray(runs, np.float32)


def ano(records):
    """ANO curve (points of the record minus the trailing initial worst) in binary32."""
    R = np.asarray(records, np.float32)
    runs, T, n = R.shape
    mn = R[:, :, n - 2].min(0)
    mx = R[:, :, n - 1].max(0)
    den = (mx - mn).astype(np.float32)
    out = np.zeros(n - 1, np.float32)
    # v - min and its quotient are binary32 operations; the running sum adds one
    # run at a time in task-major order
    live = den != 0
    q = np.zeros((runs, T, n - 1), np.float32)
    q[:, live, :] = (R[:, live, : n - 1] - mn[None, live, None]) / den[None, live, None]
    for i in range(T):
        for k in range(runs):
            out = (out + q[k, i]).astype(np.float32)
    return (out / np.float32(T * runs)).astype(np.float32)


def format_ano(curve):
    """The curve as the reference's program prints it ("v, v, ..., ")."""
    return "".join(_fmt(v) + ", " for v in curve)
