"""Online task admission: tasks that arrive in batches while earlier ones still run.

The reference's driver (proposed model/SG-EMTF.cu:1842-1958) splits its T tasks
into BATCH_NUM batches at random cut points, lets batch i arrive after a random
gap of U[0, 100) ms, starts one host thread per task as its batch arrives and
reports the mean response time (arrival -> start) and solve time (arrival ->
finished) per task.

Here a batch is one engine handle created when the batch arrives and run on its
own non-blocking stream, so batches overlap on the device with the ones already
running; nothing waits on the host except for the next arrival.  Tasks of a batch
exchange elites among themselves (a handle's transfer partners are its own
tasks).  Batch b runs with seed `batch_seed(seed, b)`, so its draws are those of a
plain Engine over the same tasks and seed: the results are bit-exact with the
oracle per batch.
"""
import time

import numpy as np

from . import hip
from .engine import Engine


def batch_seed(seed, b):
    """The Philox seed of batch b (distinct streams for tasks that share an index)."""
    return (int(seed) + int(b) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF


def arrival_schedule(n_tasks, n_batches, max_gap_ms=100.0, seed=0):
    """(cuts, arrival_ms): batch b holds tasks cuts[b]:cuts[b+1]; it arrives at arrival_ms[b].

    As the reference draws them: n_batches - 1 distinct cut points in (0, n_tasks), sorted,
    and gaps U[0, max_gap_ms) between consecutive arrivals (the first at 0)."""
    if not 1 <= n_batches <= n_tasks:
        raise ValueError("need 1 <= n_batches <= n_tasks")
    rng = np.random.default_rng(seed)
    inner = np.sort(rng.choice(np.arange(1, n_tasks), n_batches - 1, replace=False)) if n_batches > 1 else []
    cuts = np.concatenate([[0], inner, [n_tasks]]).astype(int)
    gaps = rng.uniform(0.0, max_gap_ms, n_batches)
    arrival = np.concatenate([[0.0], np.cumsum(gaps[:-1])])
    return cuts, arrival


class OnlineRun:
    """Admit batches of tasks at their arrival times and run each for `generations`.

    Each running batch keeps at most `depth` chunks of `chunk` generations queued on its
    stream; the host tops them up as they finish.  Long backlogs would otherwise sit in the
    few hardware queues that all streams share, and a new batch's first kernels would wait
    behind them."""

    def __init__(self, tasks, cuts, arrival_ms, generations, seed=0, chunk=100, depth=2, **params):
        self.tasks = tasks
        self.cuts = np.asarray(cuts, int)
        self.arrival_ms = np.asarray(arrival_ms, float)
        if len(self.cuts) != len(self.arrival_ms) + 1 or self.cuts[0] != 0 or self.cuts[-1] != len(tasks):
            raise ValueError("cuts must run from 0 to n_tasks with one batch per arrival")
        if np.any(np.diff(self.arrival_ms) < 0):
            raise ValueError("arrivals must be non-decreasing")
        if chunk < 1 or depth < 1:
            raise ValueError("chunk and depth must be >= 1")
        self.generations = int(generations)
        self.seed = seed
        self.chunk, self.depth = int(chunk), int(depth)
        # the autotuner's create-time trials free device buffers (a device-wide
        # synchronisation that would wait for the batches already running) and would be
        # timed against them: off unless asked for
        self.params = dict(params)
        self.params.setdefault("tune", 1)
        self.engines = []
        self.streams = []   # a handle keeps using its run's stream: they live as long as it does

    def run(self):
        """Returns per-task response_ms (arrival -> first generations queued, creation
        included) and solve_ms (arrival -> last generation finished)."""
        n_b = len(self.arrival_ms)
        resp = np.zeros(n_b)
        solve = np.full(n_b, np.nan)
        queued = [0] * n_b
        pending = [[] for _ in range(n_b)]
        t0 = time.perf_counter()

        def now_ms():
            return (time.perf_counter() - t0) * 1e3

        def pump():
            for b, eng in enumerate(self.engines):
                if not np.isnan(solve[b]):
                    continue
                while pending[b] and pending[b][0].query():
                    pending[b].pop(0)
                if queued[b] == self.generations and not pending[b]:
                    solve[b] = now_ms() - self.arrival_ms[b]
                    continue
                while len(pending[b]) < self.depth and queued[b] < self.generations:
                    g = min(self.chunk, self.generations - queued[b])
                    eng.run(g, stream=self.streams[b].handle)
                    ev = hip.Event()
                    ev.record(self.streams[b].handle)
                    pending[b].append(ev)
                    queued[b] += g

        for b in range(n_b):
            while now_ms() < self.arrival_ms[b]:
                pump()
                time.sleep(5e-5)
            lo, hi = self.cuts[b], self.cuts[b + 1]
            self.engines.append(Engine(self.tasks[lo:hi], seed=batch_seed(self.seed, b), **self.params))
            self.streams.append(hip.Stream())
            pump()
            resp[b] = now_ms() - self.arrival_ms[b]
        while np.isnan(solve).any():
            pump()
            time.sleep(5e-5)
        self.makespan_ms = now_ms()
        per_task = np.repeat(np.arange(n_b), np.diff(self.cuts))
        self.response_ms = resp[per_task]
        self.solve_ms = solve[per_task]
        return self.response_ms, self.solve_ms

    def best(self):
        """Best fitness of every task, in the original task order."""
        return np.concatenate([e.best()[0] for e in self.engines])

    def close(self):
        for e in self.engines:
            e.close()
        self.engines, self.streams = [], []
