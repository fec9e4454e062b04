"""Online admission (tasks start in batches while earlier ones run).

CPU: the oracle's semantics pinned -- every task admitted at 0 is the plain engine; a task
admitted later runs its own generation counter from 0 (its trajectory is that of a plain
engine over it alone when nothing transfers); tasks not yet admitted, just admitted or
finished are never transfer donors; the donor drawn is the k-th other donor in id order,
k = w mod (n - 1), as the reference draws rand() % T over every other task
(proposed model/SG-EMTF.cu:1696-1699).
GPU: the engine with online admission equals the oracle bit for bit, with transfers
across batches, batches finishing while others run, and the driver's timed admissions.
"""
import numpy as np
import pytest

from emtf import tasks as tg


def _same(a, b, ids_a, ids_b=None, tag=""):
    for ta, tb in zip(ids_a, ids_b if ids_b is not None else ids_a):
        xa, fa = a.population(ta)
        xb, fb = b.population(tb)
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), (tag, ta, tb)
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), (tag, ta, tb)


def _differs(a, b, ids):
    return any(not np.array_equal(a.population(t)[1].view(np.uint32), b.population(t)[1].view(np.uint32)) for t in ids)


PARAMS = dict(seed=5, F=0.5, CR=0.9, transfer_interval=10, transfer_count=3)


def test_all_admitted_at_zero_is_the_plain_engine(oracle_mod):
    tasks = tg.benchmark_tasks(10, 12, 32, rotate=True, seed=2)
    plain = oracle_mod.Engine(tasks, **PARAMS)
    online = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    online.admit(0, 10, 1000)
    for step in (10, 15, 5, 30):
        plain.run(step)
        online.run(step)
        _same(plain, online, range(10), tag=f"gen {plain.generation}")


def test_late_task_runs_its_own_generations(oracle_mod):
    """Without transfer a task admitted at clock 20 evolves as a plain engine over it alone
    (same task id) from generation 0."""
    tasks = tg.benchmark_tasks(6, 10, 32, rotate=True, seed=3)
    p = dict(seed=9, F=0.5, CR=0.9)
    online = oracle_mod.Engine(tasks, deferred=True, **p)
    online.admit(0, 3, 1000)
    online.run(20)
    online.admit(3, 3, 30)
    online.run(45)  # the late tasks stop after their 30 generations
    early = oracle_mod.Engine(tasks, **p)
    early.run(65)
    _same(online, early, range(3), tag="early batch")
    late = oracle_mod.Engine(tasks, **p)  # the same ids, 30 generations
    late.run(30)
    _same(online, late, range(3, 6), tag="late batch")


def test_admission_rules(oracle_mod):
    tasks = tg.benchmark_tasks(4, 6, 16, seed=1)
    e = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    e.admit(0, 2, 50)
    with pytest.raises(ValueError):
        e.admit(1, 2, 50)  # task 1 again
    e.run(5)
    with pytest.raises(ValueError):
        e.admit(2, 1, 50)  # off a transfer boundary
    e.run(5)
    e.admit(2, 2, 50)
    with pytest.raises(ValueError):
        e.admit(3, 2, 50)  # out of range


def test_new_and_pending_tasks_are_not_donors_transfer_crosses_batches(oracle_mod):
    """Batch B (tasks 8..15) admitted at clock 20 changes nothing in batch A (0..7) up to
    clock 30 (at 20 it is new, before that pending); from the boundary at 30 on, A's tasks
    draw donors from B too, and A's populations part from those of an engine without B."""
    tasks = tg.benchmark_tasks(16, 12, 32, rotate=True, seed=4)
    with_b = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    without_b = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    for e in (with_b, without_b):
        e.admit(0, 8, 1000)
        e.run(20)
    with_b.admit(8, 8, 1000)
    for e in (with_b, without_b):
        e.run(10)
    _same(with_b, without_b, range(8), tag="clock 30")
    # at 20 no one draws from B (just admitted; its own tasks receive nothing there)
    assert all(not 8 <= with_b.transfer_donor(t, 20) < 16 for t in range(16))
    for e in (with_b, without_b):
        e.run(10)
    assert _differs(with_b, without_b, range(8))
    assert any(with_b.transfer_donor(t, 30) >= 8 for t in range(8))


def test_finished_tasks_are_not_donors(oracle_mod):
    """Batch A runs 10 generations and is done at the first boundary: it is never a donor,
    so batch C evolves as in an engine where A never started."""
    tasks = tg.benchmark_tasks(16, 12, 32, rotate=True, seed=6)
    with_a = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    without_a = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    with_a.admit(0, 8, 10)
    for e in (with_a, without_a):
        e.admit(8, 8, 1000)
        e.run(40)
    _same(with_a, without_a, range(8, 16), tag="batch C")
    for t in range(16):
        assert not 0 <= with_a.transfer_donor(t, 10) < 8  # nobody draws from A (A itself: -1)


def test_donor_is_the_kth_other_donor(oracle_mod):
    """The donor at a boundary is the k-th, in id order, of the other tasks that ran up to
    it and run on after it, k = w mod (n - 1) with w the task's own transfer draw."""
    tasks = tg.benchmark_tasks(24, 6, 16, seed=7)
    e = oracle_mod.Engine(tasks, deferred=True, **PARAMS)
    life = {}  # task -> (start, end)
    e.admit(0, 6, 30)
    life.update({t: (0, 30) for t in range(0, 6)})
    e.admit(6, 6, 1000)
    life.update({t: (0, 1000) for t in range(6, 12)})
    e.run(20)
    e.admit(12, 5, 1000)
    life.update({t: (20, 1020) for t in range(12, 17)})
    e.run(20)  # clock 40: tasks 0..5 done at 30, 12..16 running since 20, 17..23 pending
    for clock in (30, 40, 50):
        elig = [t for t, (a, b) in life.items() if a < clock < b]
        for t in range(24):
            got = e.transfer_donor(t, clock)
            if t not in elig:
                assert got == -1
                continue
            w = int(oracle_mod.draw_block(PARAMS["seed"], t, clock - life[t][0], 7, 0, 0, 0)[0])
            others = [u for u in elig if u != t]
            assert got == others[w % len(others)], (clock, t)


# ---------------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_online_admission_bit_exact(oracle_mod):
    """Batches admitted at chosen clocks, each for its own number of generations, runs of
    every length (epochs cut at transfer boundaries and where a batch finishes), LDS and wide
    kernels, islands: the engine equals the oracle after every step, tasks not admitted yet
    included (zero rows, NaN fitness)."""
    import emtf
    dims = [12, 50, 150, 40, 20, 64, 100, 33] * 3
    pops = [32, 64, 96, 128, 64, 32, 160, 64] * 3
    tasks = tg.benchmark_tasks(24, 0, 0, rotate=True, seed=8, dims=dims, pops=pops)
    params = dict(seed=13, F=0.5, CR=0.9, n_islands=2, exchange_interval=4, migrate_interval=5, migrate_count=2,
                  transfer_interval=10, transfer_count=3)
    print("case: online admission, 24 tasks in 3 batches", flush=True)
    gpu = emtf.Engine(tasks, admission=emtf.ADMIT_ONLINE, **params)
    orc = oracle_mod.Engine(tasks, deferred=True, **params)
    plan = [("admit", 0, 8, 60), ("run", 10), ("run", 10), ("admit", 8, 10, 100), ("run", 15), ("run", 5),
            ("admit", 18, 6, 40), ("run", 40), ("run", 7), ("run", 33), ("run", 30)]
    for step in plan:
        for e in (gpu, orc):
            if step[0] == "admit":
                e.admit(*step[1:])
            else:
                e.run(step[1])
        if step[0] == "run":
            _same(gpu, orc, range(24), tag=f"clock {orc.generation}")
    with pytest.raises(emtf.EmtfError):
        gpu.admit(0, 1, 10)  # admitted before


@pytest.mark.gpu
def test_gpu_online_run_replays_on_the_oracle(oracle_mod):
    """The timed driver: batches arrive at U[0, 15) ms gaps and are admitted at the clock the
    queued work has reached; replaying the recorded admissions on the oracle gives every
    task's population bit for bit; response <= solve for every task."""
    from emtf import online
    tasks = tg.uneven_tasks(30, seed=6, dmax=120, pmax=128)
    cuts, arr = online.arrival_schedule(len(tasks), 4, max_gap_ms=15.0, seed=1)
    params = dict(transfer_interval=5, transfer_count=2)
    print("case: timed online run, 30 tasks in 4 batches", flush=True)
    run = online.OnlineRun(tasks, cuts, arr, 40, seed=11, chunk=10, **params)
    resp, solve = run.run()
    assert np.all(resp >= 0) and np.all(solve >= resp)
    assert np.all(np.diff(run.admitted_at) >= 0) and np.all(run.admitted_at % 5 == 0)
    orc = oracle_mod.Engine(tasks, deferred=True, seed=11, **params)
    for b in range(len(arr)):
        orc.run(int(run.admitted_at[b]) - orc.generation)
        orc.admit(cuts[b], cuts[b + 1] - cuts[b], 40)
    orc.run(run.clock - orc.generation)
    _same(run.engine, orc, range(len(tasks)), tag="timed run")
    run.close()
