"""Gradient polish of elites (the adjoint driving Armijo steps between generations).

CPU: the oracle's polish against an unpolished run on the same draws -- only the
top-k rows change, none gets worse, stored fitness equals a fresh evaluation,
and on the Sphere it closes most of the gap.  GPU: the engine with polish
enabled (through the C ABI) is bit-exact with the oracle, also across transfer
boundaries where the polished elites are the ones published.
"""
import numpy as np
import pytest

import emtf
from emtf import tasks as tg


def _rows(eng, t):
    x, f = eng.population(t)
    return x.copy(), f.copy()


def test_oracle_polish_changes_only_elites_and_never_worsens(oracle_mod):
    tasks = tg.benchmark_tasks(8, 20, 32, rotate=True, seed=21)
    base = dict(seed=4, F=0.5, CR=0.9)
    plain = oracle_mod.Engine(tasks, **base)
    pol = oracle_mod.Engine(tasks, **base, polish_interval=5, polish_steps=4, polish_topk=4)
    plain.run(5)
    pol.run(5)
    gains = []
    for t, task in enumerate(tasks):
        x0, f0 = _rows(plain, t)
        x1, f1 = _rows(pol, t)
        key = np.where(np.isnan(f0), np.inf, f0)
        elite = np.lexsort((np.arange(len(f0)), key))[:4]
        others = np.setdiff1d(np.arange(len(f0)), elite)
        assert np.array_equal(x0[others].view(np.uint32), x1[others].view(np.uint32)), t
        assert np.array_equal(f0[others].view(np.uint32), f1[others].view(np.uint32)), t
        assert np.all(f1[elite] <= f0[elite]), (t, f0[elite], f1[elite])
        fresh = oracle_mod.eval32(task, x1[elite])
        assert np.array_equal(fresh.view(np.uint32), f1[elite].view(np.uint32)), t
        gains.append(float(f0[elite].min() - f1[elite].min()))
    # Sphere (task 0): Polyak steps on a convex quadratic land near the optimum
    x0, f0 = _rows(plain, 0)
    _, f1 = _rows(pol, 0)
    assert f1.min() < 1e-2 * f0.min()
    assert sum(g > 0 for g in gains) >= 4


def test_oracle_polish_off_is_identity(oracle_mod):
    tasks = tg.benchmark_tasks(4, 10, 32, rotate=False, seed=2)
    a = oracle_mod.Engine(tasks, seed=1)
    b = oracle_mod.Engine(tasks, seed=1, polish_interval=5, polish_steps=0, polish_topk=4)
    a.run(7)
    b.run(7)
    for t in range(4):
        assert np.array_equal(_rows(a, t)[0].view(np.uint32), _rows(b, t)[0].view(np.uint32))


def test_polish_rejects_negative_knobs():
    import os
    if not os.path.exists(emtf.LIB_PATH):
        pytest.skip("libemtf.so not built")
    tasks = tg.benchmark_tasks(2, 10, 32, rotate=False, seed=2)
    with pytest.raises(emtf.EmtfError, match="polish"):
        emtf.Engine(tasks, polish_interval=-1, polish_steps=1, polish_topk=1)


def _compare(gpu, orc, tag):
    for t in range(gpu.n_tasks):
        xg, fg = gpu.population(t)
        xo, fo = orc.population(t)
        assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32)), f"{tag}: task {t} fitness differs"
        assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), f"{tag}: task {t} population differs"


@pytest.mark.gpu
@pytest.mark.parametrize("dim,pop,topk,steps", [(50, 64, 4, 3), (20, 128, 40, 2), (128, 128, 64, 2), (7, 32, 32, 5),
                                                 (50, 64, 3, 4), (33, 96, 12, 3), (100, 128, 8, 5), (50, 128, 1, 5),
                                                 (64, 256, 200, 1)])
def test_gpu_polish_bit_exact(oracle_mod, dim, pop, topk, steps):
    tasks = tg.benchmark_tasks(8, dim, pop, rotate=True, seed=30 + dim)
    params = dict(seed=12, F=0.5, CR=0.9, transfer_interval=10, transfer_count=2, polish_interval=5,
                  polish_steps=steps, polish_topk=topk)
    print(f"case: polish D={dim} pop={pop} topk={topk} steps={steps}", flush=True)
    gpu = emtf.Engine(tasks, **params)
    orc = oracle_mod.Engine(tasks, **params)
    for step in (5, 7, 8):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"polish gen {orc.generation}")


@pytest.mark.gpu
def test_gpu_polish_across_buckets(oracle_mod):
    """Polish with tasks of several row geometries (one LDS kernel bucket each, on their
    own streams), with transfer publishing the polished elites across buckets."""
    dims = [12, 40, 70, 100, 12, 40, 70, 100]
    tasks = tg.benchmark_tasks(8, 0, 64, rotate=True, seed=17, dims=dims, pops=[64, 96, 64, 128, 32, 64, 96, 64])
    params = dict(seed=2, F=0.5, CR=0.9, transfer_interval=5, transfer_count=2, polish_interval=5,
                  polish_steps=3, polish_topk=4)
    gpu = emtf.Engine(tasks, **params)
    orc = oracle_mod.Engine(tasks, **params)
    for step in (5, 10, 3):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"polish buckets gen {orc.generation}")
