"""Online task admission: tasks that arrive in batches while earlier ones still run.

The reference's driver (proposed model/SG-EMTF.cu:1842-1958) splits its T tasks
into BATCH_NUM batches at random cut points, lets batch i arrive after a random
gap of U[0, 100) ms, starts one host thread per task as its batch arrives and
reports the mean response time (arrival -> start) and solve time (arrival ->
finished) per task; a task's transfer partner is drawn among every task still
running (:1682-1699).

Here all tasks live in one engine handle created with online admission: the
handle's generation counter is a clock that advances in chunks queued on one
stream, and a batch that has arrived is admitted (emtf_admit) at the clock the
queued work has reached, so it waits behind at most `depth` chunks.  Its tasks
then run their own generations 0.. and exchange elites with every task running
at a transfer boundary, whichever batch it came with.  The clock at which each
batch was admitted is recorded (`admitted_at`): replaying the same admissions on
the oracle reproduces every population bit for bit.
"""
import time

import numpy as np

from . import hip
from .engine import ADMIT_ONLINE, Engine


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

    The clock advances in chunks of `chunk` generations (a multiple of the transfer
    interval, where a batch may be admitted), at most `depth` of them queued ahead."""

    def __init__(self, tasks, cuts, arrival_ms, generations, seed=0, chunk=20, depth=2, **params):
        self.tasks = tasks
        self.cuts = np.asarray(cuts, int)
        self.arrival_ms = np.asarray(arrival_ms, float)
        if len(self.cuts) != len(self.arrival_ms) + 1 or self.cuts[0] != 0 or self.cuts[-1] != len(tasks):
            raise ValueError("cuts must run from 0 to n_tasks with one batch per arrival")
        if np.any(np.diff(self.arrival_ms) < 0):
            raise ValueError("arrivals must be non-decreasing")
        if chunk < 1 or depth < 1:
            raise ValueError("chunk and depth must be >= 1")
        ti = int(params.get("transfer_interval", 0))
        if ti > 0 and chunk % ti:
            raise ValueError("chunk must be a multiple of the transfer interval")
        self.generations = int(generations)
        self.chunk, self.depth = int(chunk), int(depth)
        self.params = dict(params, seed=seed)
        self.engine = Engine(tasks, admission=ADMIT_ONLINE, **self.params)
        self.stream = hip.Stream()
        self.admitted_at = np.full(len(self.arrival_ms), -1, int)  # clock of each admission

    def run(self):
        """Returns per-task response_ms (arrival -> admitted, its initial population queued)
        and solve_ms (arrival -> its last generation finished)."""
        n_b = len(self.arrival_ms)
        resp = np.zeros(n_b)
        solve = np.full(n_b, np.nan)
        eng, st = self.engine, self.stream.handle
        clock = 0          # the clock the queued work reaches
        queued = []        # (event, clock at its end)
        nxt = 0            # next batch to arrive
        t0 = time.perf_counter()

        def now_ms():
            return (time.perf_counter() - t0) * 1e3

        while np.isnan(solve).any():
            while nxt < n_b and now_ms() >= self.arrival_ms[nxt]:
                lo, hi = self.cuts[nxt], self.cuts[nxt + 1]
                eng.admit(lo, hi - lo, self.generations, stream=st)
                self.admitted_at[nxt] = clock
                resp[nxt] = now_ms() - self.arrival_ms[nxt]
                nxt += 1
            while queued and queued[0][0].query():
                done = queued.pop(0)[1]
                for b in range(nxt):
                    if np.isnan(solve[b]) and self.admitted_at[b] + self.generations <= done:
                        solve[b] = now_ms() - self.arrival_ms[b]
            running = any(self.admitted_at[b] + self.generations > clock for b in range(nxt))
            if running and len(queued) < self.depth:
                eng.run(self.chunk, stream=st)
                ev = hip.Event()
                ev.record(st)
                clock += self.chunk
                queued.append((ev, clock))
            else:
                time.sleep(5e-5)
        self.clock = clock
        self.makespan_ms = now_ms()
        per_task = np.repeat(np.arange(n_b), np.diff(self.cuts))
        self.response_ms = resp[per_task]
        self.solve_ms = solve[per_task]
        return self.response_ms, self.solve_ms

    def best(self):
        """Best fitness of every task, in the original task order."""
        return self.engine.best()[0]

    def close(self):
        self.engine.close()
