"""CVRP GA: the oracle pinned by brute force and operator properties (CPU), and the
HIP kernel bit-exact with the oracle (GPU)."""
import itertools

import numpy as np
import pytest

from emtf.cvrp import random_instances


def _small(n, seed, cap=15, pop=16):
    rng = np.random.default_rng(seed)
    xy = np.vstack([[50, 50], rng.integers(0, 101, (n, 2))]).astype(np.int32)
    dem = np.concatenate([[0], rng.integers(1, 11, n)]).astype(np.int32)
    return dict(n=n, capacity=cap, pop=pop, xy=xy, demand=dem)


def test_nint_distances(oracle_mod):
    ins = _small(12, 1)
    d = oracle_mod.cvrp_dist_matrix(ins["xy"])
    xy = ins["xy"].astype(np.float64)
    ref = np.floor(np.sqrt(((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1)) + 0.5)
    assert np.array_equal(d, ref.astype(np.uint8))
    assert d.max() <= 141


def _brute_split(tour, d, dem, cap):
    n = len(tour)
    best = None
    for mask in range(1 << (n - 1)):
        cuts = [0] + [k + 1 for k in range(n - 1) if mask >> k & 1] + [n]
        cost, ok = 0, True
        for s, e in zip(cuts[:-1], cuts[1:]):
            r = tour[s:e]
            if dem[r].sum() > cap:
                ok = False
                break
            cost += d[0, r[0]] + sum(d[r[k], r[k + 1]] for k in range(len(r) - 1)) + d[r[-1], 0]
        if ok and (best is None or cost < best):
            best = cost
    return best


@pytest.mark.parametrize("seed", range(6))
def test_split_is_the_optimal_partition(oracle_mod, seed):
    n = 9
    ins = _small(n, seed, cap=14)
    d = oracle_mod.cvrp_dist_matrix(ins["xy"]).astype(np.int64)
    dem = ins["demand"]
    rng = np.random.default_rng(100 + seed)
    for _ in range(5):
        tour = rng.permutation(np.arange(1, n + 1)).astype(np.uint8)
        got = oracle_mod.cvrp_split(tour, d.astype(np.uint8), dem.astype(np.uint8), ins["capacity"])
        assert got == _brute_split(tour.astype(np.int64), d, dem, ins["capacity"])


def test_order_crossover_and_mutations(oracle_mod):
    rng = np.random.default_rng(3)
    n = 20
    for trial in range(200):
        p1 = rng.permutation(np.arange(1, n + 1)).astype(np.uint8)
        p2 = rng.permutation(np.arange(1, n + 1)).astype(np.uint8)
        a, b = sorted(rng.integers(0, n, 2))
        c = oracle_mod.cvrp_child(p1, p2, int(a), int(b), 0, 0, 0)
        assert sorted(c) == list(range(1, n + 1))
        assert np.array_equal(c[a:b + 1], p1[a:b + 1])
        # the other positions, read cyclically from b+1, hold p2's remaining genes in p2's cyclic order from b+1
        pos = [(b + 1 + k) % n for k in range(n - (b - a + 1))]
        rest = [g for g in np.roll(p2, -((b + 1) % n)) if g not in set(p1[a:b + 1])]
        assert [int(c[p]) for p in pos] == [int(g) for g in rest]
        x, y = (int(v) for v in rng.integers(0, n, 2))
        sw = oracle_mod.cvrp_child(p1, p2, int(a), int(b), 1, x, y)
        c2 = c.copy()
        c2[x], c2[y] = c2[y], c2[x]
        assert np.array_equal(sw, c2)
        tw = oracle_mod.cvrp_child(p1, p2, int(a), int(b), 2, x, y)
        lo, hi = min(x, y), max(x, y)
        c3 = c.copy()
        c3[lo:hi + 1] = c3[lo:hi + 1][::-1]
        assert np.array_equal(tw, c3)


def test_engine_invariants(oracle_mod):
    inst = [_small(n, 10 + n, cap=30, pop=32) for n in (7, 15, 40)]
    eng = oracle_mod.CvrpEngine(inst, p_swap=0.3, p_2opt=0.4, seed=5)
    prev = [eng.population(i)[1].copy() for i in range(3)]
    for _ in range(4):
        eng.run(5)
        for i, ins in enumerate(inst):
            t, c = eng.population(i)
            d = oracle_mod.cvrp_dist_matrix(ins["xy"])
            for k in range(ins["pop"]):
                assert sorted(t[k]) == list(range(1, ins["n"] + 1))
                assert c[k] == oracle_mod.cvrp_split(t[k], d, ins["demand"].astype(np.uint8), ins["capacity"])
            assert np.all(c <= prev[i])
            prev[i] = c.copy()


def test_ga_finds_brute_force_optimum(oracle_mod):
    n = 6
    ins = _small(n, 42, cap=12, pop=16)
    d = oracle_mod.cvrp_dist_matrix(ins["xy"]).astype(np.int64)
    opt = min(_brute_split(np.array(p), d, ins["demand"], ins["capacity"])
              for p in itertools.permutations(range(1, n + 1)))
    eng = oracle_mod.CvrpEngine([ins], p_swap=0.3, p_2opt=0.3, seed=1)
    eng.run(200)
    assert eng.population(0)[1].min() == opt


@pytest.mark.gpu
@pytest.mark.parametrize("n_islands", [1, 4])
def test_gpu_cvrp_bit_exact(oracle_mod, n_islands):
    from emtf.cvrp import CvrpEngine
    inst = random_instances(12, seed=3, nmin=20, nmax=200, pop=64)
    inst += [_small(5, 1, cap=12, pop=64), _small(255, 2, cap=50, pop=64)]
    params = dict(p_swap=0.3, p_2opt=0.3, n_islands=n_islands, migrate_interval=5 if n_islands > 1 else 0,
                  migrate_count=2 if n_islands > 1 else 0, seed=11)
    print("case: cvrp", len(inst), "instances, islands", n_islands, flush=True)
    gpu = CvrpEngine(inst, **params)
    orc = oracle_mod.CvrpEngine(inst, **params)
    for step in (0, 7, 13):
        if step:
            gpu.run(step)
            orc.run(step)
        for i in range(len(inst)):
            tg, cg = gpu.population(i)
            to, co = orc.population(i)
            assert np.array_equal(cg, co), (step, i, cg[:4], co[:4])
            assert np.array_equal(tg, to), (step, i)


@pytest.mark.gpu
def test_gpu_cvrp_bit_exact_32bit_labels(oracle_mod):
    """Instances whose split ring values can exceed 2^16 (corner depot, spread coordinates)
    run with 32 + 32-bit ring entries; results stay bit-exact."""
    from emtf.cvrp import CvrpEngine
    rng = np.random.default_rng(5)
    inst = []
    for n in (255, 240, 60):
        xy = np.vstack([[0, 0], rng.integers(0, 181, (n, 2))]).astype(np.int32)
        dem = np.concatenate([[0], rng.integers(1, 11, n)]).astype(np.int32)
        inst.append(dict(n=n, capacity=50, pop=64, xy=xy, demand=dem))
    d0 = oracle_mod.cvrp_dist_matrix(inst[0]["xy"])[0]
    assert 2 * int(d0[1:].astype(np.int64).sum()) >= 65536
    params = dict(p_swap=0.3, p_2opt=0.3, n_islands=2, migrate_interval=4, migrate_count=2, seed=3)
    gpu = CvrpEngine(inst, **params)
    orc = oracle_mod.CvrpEngine(inst, **params)
    gpu.run(9)
    orc.run(9)
    for i in range(len(inst)):
        tg, cg = gpu.population(i)
        to, co = orc.population(i)
        assert np.array_equal(cg, co), i
        assert np.array_equal(tg, to), i


@pytest.mark.gpu
def test_gpu_cvrp_long_routes(oracle_mod):
    """Unit demands and capacity 60: routes (the split's windows) of up to 60 customers,
    past the split's fixed 16-entry scan; bit-exact with the oracle."""
    from emtf.cvrp import CvrpEngine
    rng = np.random.default_rng(8)
    inst = []
    for n in (255, 130, 61, 7):
        xy = np.vstack([[50, 50], rng.integers(0, 101, (n, 2))]).astype(np.int32)
        inst.append(dict(n=n, capacity=60, pop=64, xy=xy, demand=np.concatenate([[0], np.ones(n)]).astype(np.int32)))
    params = dict(p_swap=0.3, p_2opt=0.3, n_islands=2, migrate_interval=3, migrate_count=2, seed=4)
    print("case: cvrp long routes", flush=True)
    gpu = CvrpEngine(inst, **params)
    orc = oracle_mod.CvrpEngine(inst, **params)
    gpu.run(11)
    orc.run(11)
    for i in range(len(inst)):
        tg, cg = gpu.population(i)
        to, co = orc.population(i)
        assert np.array_equal(cg, co), i
        assert np.array_equal(tg, to), i


@pytest.mark.gpu
def test_gpu_cvrp_full_size_properties(oracle_mod):
    """The CVRP benchmark batch at its real size (512 instances, n ~ U{50..200}, pop 128,
    4 islands): over all 5000 generations each instance's best cost never increases, every
    tour stays a permutation of the customers, and every stored cost equals the oracle's
    optimal split of the stored tour (sampled instances)."""
    from emtf.cvrp import CvrpEngine
    inst = random_instances(512, seed=0)
    gpu = CvrpEngine(inst, p_swap=0.3, p_2opt=0.3, n_islands=4, migrate_interval=25, migrate_count=2, seed=0)
    prev = gpu.best()[0]
    for _ in range(10):
        gpu.run(500)
        cur = gpu.best()[0]
        assert np.all(cur <= prev)
        prev = cur
    for i in range(0, 512, 37):
        ins = inst[i]
        dist = oracle_mod.cvrp_dist_matrix(ins["xy"])
        dem = ins["demand"].astype(np.uint8)
        tours, costs = gpu.population(i)
        for k in range(0, ins["pop"], 9):
            assert sorted(tours[k].tolist()) == list(range(1, ins["n"] + 1))
            assert costs[k] == oracle_mod.cvrp_split(tours[k], dist, dem, ins["capacity"]), (i, k)
