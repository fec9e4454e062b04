"""Convergence records, ANO and the FP64 error metric.

ANO is pinned by the reference's own data: its T = 100 record file (18 runs x
100 tasks) and the curve its evaluation program printed for it must come out
digit for digit.  error64 is pinned against the oracle's independent FP64
evaluator and against the known minima.
"""
import os

import numpy as np
import pytest

from emtf import convergence as cv
from emtf import tasks as tg

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_ano_reproduces_reference_curve():
    R = np.load(os.path.join(GOLD, "ano_T100_records.npz"))["records"]
    assert R.shape == (18, 100, 22)
    with open(os.path.join(GOLD, "ano_T100_expected.txt")) as fh:
        expected = fh.read().strip()
    assert cv.format_ano(cv.ano(R)).strip() == expected


def test_ano_constant_task_contributes_zero_and_range():
    R = np.ones((2, 3, 5), np.float32)
    R[:, 1, :] = np.array([10, 5, 2, 1, 20], np.float32)
    curve = cv.ano(R)
    # task 1: (v - 1) / (20 - 1); tasks 0 and 2 are constant
    ref = np.array([9, 4, 1, 0], np.float32) / np.float32(19) / np.float32(3)
    assert np.allclose(curve, ref, rtol=1e-6)
    assert curve[-1] == 0


def test_record_roundtrip(tmp_path):
    R = np.load(os.path.join(GOLD, "ano_T100_records.npz"))["records"][:3]
    path = str(tmp_path / "rec.txt")
    for run in R:
        cv.write_record(path, run)
    back = cv.read_record(path)
    assert np.array_equal(back.view(np.uint32), R.view(np.uint32))
    lines = open(path).read().split("\n")
    assert lines[0].endswith(" ") and lines[100] == ""


def test_error64_matches_oracle_eval64(oracle_mod):
    tasks = tg.benchmark_tasks(16, 30, 32, rotate=True, seed=5)
    rng = np.random.default_rng(1)
    for t in tasks:
        x = (t["shift"][None, :] + rng.normal(0, 2.0, (8, 30))).astype(np.float32)
        a = cv.objective64(t, x)
        b = oracle_mod.eval64(t, x.astype(np.float64))
        assert np.allclose(a, b, rtol=1e-9, atol=1e-9), (t["func"], a, b)


def test_error64_at_minimum():
    tasks = tg.benchmark_tasks(8, 20, 32, rotate=True, seed=9)
    xb = np.zeros((8, 24))
    for i, t in enumerate(tasks):
        z = np.ones(20) if t["func"] == 1 else np.zeros(20)
        xb[i, :20] = t["shift"].astype(np.float64) + np.asarray(t["rot"], np.float64).T @ z
    err = cv.error64(tasks, xb)
    assert np.isnan(err[5])  # Schwefel: no known minimum in the box
    assert np.all(np.abs(np.delete(err, 5)) < 1e-9), err


@pytest.mark.gpu
def test_recorder_on_engine(tmp_path):
    import emtf
    tasks = tg.benchmark_tasks(16, 20, 64, rotate=True, seed=3)
    eng = emtf.Engine(tasks, seed=2, transfer_interval=10, transfer_count=2)
    rec = cv.ConvergenceRecorder(eng, 10)
    rec.run(100)
    rows = rec.rows()
    assert rows.shape == (16, 12)
    assert np.all(np.diff(rows[:, :-1], axis=1) <= 0)  # elitist: the best never gets worse
    assert np.all(rows[:, -1] >= rows[:, 0])
    path = str(tmp_path / "rec.txt")
    cv.write_record(path, rows)
    curve = cv.ano(cv.read_record(path))
    assert curve[0] == pytest.approx(np.mean((rows[:, 0] - rows[:, -2]) / (rows[:, -1] - rows[:, -2])), rel=1e-4)


def test_para_roundtrip(tmp_path):
    tasks = tg.reference_tasks(3, 5, 32, seed=1) + tg.benchmark_tasks(2, 4, 32, rotate=True, seed=2)
    p, r = str(tmp_path / "para.txt"), str(tmp_path / "rot.npz")
    cv.write_para(p, tasks, rotations_path=r)
    M, b, ids = cv.read_para(p)
    assert ids == [t["func"] for t in tasks]
    for t, task in enumerate(tasks):
        o = task["shift"]
        assert np.allclose(b[t], -o, rtol=1e-5, atol=1e-5)
        want = task["diag"] if task["xform"] == 2 else np.ones(task["dim"])
        assert np.allclose(M[t], want, rtol=1e-5)
    rots = np.load(r)
    assert np.array_equal(rots["task4"], tasks[4]["rot"])
    with pytest.raises(ValueError):
        cv.write_para(p, tasks)


@pytest.mark.gpu
def test_elite_export_matches_population_order():
    import emtf
    tasks = tg.uneven_tasks(12, seed=3)
    eng = emtf.Engine(tasks, seed=5, transfer_interval=4, transfer_count=2)
    eng.run(9)
    k = 16
    f, x = eng.elites(k)
    for t, task in enumerate(tasks):
        xp, fp = eng.population(t)
        key = np.where(np.isnan(fp), np.inf, fp)
        order = np.lexsort((np.arange(len(fp)), key))[:k]
        assert np.array_equal(f[t].view(np.uint32), fp[order].view(np.uint32))
        D = task["dim"]
        assert np.array_equal(x[t, :, :D].view(np.uint32), xp[order].view(np.uint32))
        assert not x[t, :, D:].any()
    fb, xb = eng.best()
    assert np.array_equal(fb.view(np.uint32), f[:, 0].view(np.uint32))


@pytest.mark.gpu
def test_device_trace_matches_oracle(oracle_mod):
    """emtf_trace_start / emtf_trace record on the device what a host loop over the oracle
    records (best every 7 generations, initial worst last), with runs whose lengths do not
    line up with the interval; the traced run's populations stay bit-exact."""
    import emtf
    tasks = tg.uneven_tasks(10, seed=4)
    params = dict(seed=8, transfer_interval=5, transfer_count=2)
    eng = emtf.Engine(tasks, **params)
    orc = oracle_mod.Engine(tasks, **params)
    eng.run(3)
    orc.run(3)
    eng.trace_start(7, 6)
    want_best = [orc.best()[0]]
    worst = np.array([np.nanmax(orc.population(t)[1]) for t in range(len(tasks))], np.float32)
    for step in (5, 13, 1, 30):
        eng.run(step)
    for _ in range(6 - 1):
        orc.run(7)
        want_best.append(orc.best()[0])
    orc.run(3 + 49 - orc.generation)
    rows = eng.trace()
    assert rows.shape == (len(tasks), 7)
    want = np.concatenate([np.stack(want_best, 1), worst[:, None]], 1)
    assert np.array_equal(rows.view(np.uint32), want.view(np.uint32))
    for t in range(len(tasks)):
        assert np.array_equal(eng.population(t)[0].view(np.uint32), orc.population(t)[0].view(np.uint32))
    eng.trace_start(0, 0)
    eng.run(10)
    assert eng.trace().shape == (len(tasks), 7)


def test_arrival_schedule():
    from emtf import online
    cuts, arr = online.arrival_schedule(100, 6, max_gap_ms=50.0, seed=3)
    assert cuts[0] == 0 and cuts[-1] == 100 and len(cuts) == 7 and np.all(np.diff(cuts) > 0)
    assert arr[0] == 0.0 and len(arr) == 6 and np.all(np.diff(arr) >= 0) and np.all(np.diff(arr) < 50.0)
    assert online.arrival_schedule(5, 1)[0].tolist() == [0, 5]
    with pytest.raises(ValueError):
        online.arrival_schedule(3, 4)
