"""Pins the oracle's population-level phases: the cross-island exchange, the island
migration (continuous and CVRP) and the inter-task transfer.

Three kinds of check, none of which reads the oracle's own code:
  * the "islands" section of tests/golden/toy_generation.json, written by an independent
    numpy restatement of the reference (tests/golden/make_toy_generation.py): the oracle's
    phases one at a time and its whole run reproduce it bit for bit, and so does the GPU;
  * properties that the reference's definitions fix: the exchange permutes the rows, one
    rotation of the islands per slot (proposed model/SG-EMTF.cu:1227, select_interval
    without a block term); migration moves exactly the source's rank-j member onto the
    destination's rank m-k+j member iff strictly better (SG-EMTF_DE.cu:1127-1163, genes and
    fitness kept together); transfer maps a donor gene v to lo1 + (v - lo2) / r2 * r1 (so
    the box corners map onto the box corners exactly, SG-EMTF.cu:1555) and replaces own
    rank pop-k+j (SG-EMTF.cu:1568-1589);
  * planted mistakes: the oracle built with -DOR_MUTANT=k (a wrong rank, a non-strict
    comparison, an offset per island, swapped box ranges, a wrong target; for online
    admission a finished or just-admitted task taken as donor, the donor list skipping
    the task's id instead of its position, draws addressed by the clock) fails the
    property check meant to catch it.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import test_online

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
f32 = np.float32
P_MIG, P_XFER = 7, 8


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _arr(v, shape=None):
    a = np.array(v, np.uint32).view(np.float32)
    return a.reshape(shape) if shape else a


def _islands():
    d = json.load(open(os.path.join(GOLD, "toy_generation.json")))["islands"]
    P = d["pop"]
    tasks = [dict(func=0, dim=t["dim"], pop=P, xform=0, lo=t["lo"], hi=t["hi"], shift=_arr(t["shift"]))
             for t in d["tasks"]]
    q = d["params"]
    params = dict(seed=d["seed"], algo=0, F=q["F"], CR=q["CR"], n_islands=q["n_islands"], exchange_interval=1,
                  migrate_interval=1, migrate_count=q["migrate_count"], transfer_interval=1,
                  transfer_count=q["transfer_count"], transfer_cr=q["transfer_cr"])

    def state(rec, t):
        return _arr(rec["x"], (P, tasks[t]["dim"])), _arr(rec["fitness"])
    return d, tasks, params, state


def _same(a, b):
    return np.array_equal(_bits(a[0]), _bits(b[0])) and np.array_equal(_bits(a[1]), _bits(b[1]))


# ------------------------------------------------------------------ fixture: oracle
def test_islands_fixture_init(oracle_mod):
    d, tasks, params, state = _islands()
    eng = oracle_mod.Engine(tasks, **params)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["init"][t], t)), t


def test_islands_fixture_phases(oracle_mod):
    """Each phase of the oracle, started from the fixture's state before it, gives the
    fixture's state after it."""
    d, tasks, params, state = _islands()
    eng = oracle_mod.Engine(tasks, **params)
    prev = [state(d["init"][t], t) for t in range(len(tasks))]
    for g in d["generations"]:
        gen = g["generation"]
        if "transfer" in g:
            for t in range(len(tasks)):
                eng.set_population(t, *prev[t])
            eng.phase_transfer(gen - 1)
            prev = [state(g["after_transfer"][t], t) for t in range(len(tasks))]
            for t in range(len(tasks)):
                assert _same(eng.population(t), prev[t]), ("transfer", gen, t)
        for t, rec in enumerate(g["tasks"]):
            eng.set_population(t, *prev[t])
            eng.phase_exchange(t, gen)
            assert _same(eng.population(t), state(rec["exchange"], t)), ("exchange", gen, t)
            eng.set_population(t, *state(rec["de"], t))
            eng.phase_migrate(t)
            assert _same(eng.population(t), state(rec["migration"], t)), ("migration", gen, t)
            prev[t] = state(rec["migration"], t)


def test_islands_fixture_run(oracle_mod):
    """The oracle's whole run (exchange, DE, selection, migration, transfer) equals the fixture."""
    d, tasks, params, state = _islands()
    eng = oracle_mod.Engine(tasks, **params)
    eng.run(1)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["generations"][0]["tasks"][t]["migration"], t)), t
    eng.run(len(d["generations"]) - 1)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["final"][t], t)), t


@pytest.mark.gpu
def test_islands_fixture_gpu():
    """The GPU engine reproduces the fixture: generation 1, then the transfer and generation 2."""
    import emtf
    d, tasks, params, state = _islands()
    eng = emtf.Engine(tasks, **params)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["init"][t], t)), ("init", t)
    eng.run(1)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["generations"][0]["tasks"][t]["migration"], t)), ("gen 1", t)
    eng.run(len(d["generations"]) - 1)
    for t in range(len(tasks)):
        assert _same(eng.population(t), state(d["final"][t], t)), ("final", t)


# ------------------------------------------------------------- property checks
def _rank(f):
    key = np.where(np.isnan(f), np.inf, f)
    return list(np.lexsort((np.arange(len(f)), key)))


def check_exchange(orc):
    """Exchange permutes the task's rows; at every slot k the islands are rotated by one
    offset (the row at slot k of island q comes from island (q + off_k) mod n)."""
    rs = np.random.default_rng(3)
    n, m, D = 4, 8, 5
    task = dict(func=0, dim=D, pop=n * m, lo=-10.0, hi=10.0)
    eng = orc.Engine([task], seed=77, n_islands=n, exchange_interval=1)
    for gen in range(1, 9):
        X = rs.uniform(-10, 10, (n * m, D)).astype(f32)
        f = rs.uniform(0, 1, n * m).astype(f32)
        eng.set_population(0, X, f)
        eng.phase_exchange(0, gen)
        X1, f1 = eng.population(0)
        where = {X[i].tobytes(): i for i in range(n * m)}
        src = [where.get(X1[i].tobytes(), -1) for i in range(n * m)]
        assert sorted(src) == list(range(n * m)), f"gen {gen}: the exchange lost or duplicated rows"
        assert np.array_equal(_bits(f1), _bits(f[src])), f"gen {gen}: fitness left its row"
        for k in range(m):
            shifts = {(src[q * m + k] // m - q) % n for q in range(n)}
            assert len(shifts) == 1, f"gen {gen} slot {k}: not one rotation of the islands"


def check_migration(orc):
    """Island q's rank-j member replaces island q+1's rank m-k+j member iff strictly better
    (ranks by (fitness, index) before any move); nothing else changes."""
    rs = np.random.default_rng(5)
    n, m, k, D = 4, 8, 3, 4
    task = dict(func=0, dim=D, pop=n * m, lo=-10.0, hi=10.0)
    eng = orc.Engine([task], seed=5, n_islands=n, migrate_interval=1, migrate_count=k)
    for trial in range(8):
        X = rs.uniform(-10, 10, (n * m, D)).astype(f32)
        f = rs.integers(0, 3, n * m).astype(f32)  # many ties: 'strictly better' matters
        eng.set_population(0, X, f)
        eng.phase_migrate(0)
        X1, f1 = eng.population(0)
        ex, ef = X.copy(), f.copy()
        ords = [[q * m + r for r in _rank(f[q * m:(q + 1) * m])] for q in range(n)]
        for q in range(n):
            p = (q + 1) % n
            for j in range(k):
                s, t = ords[q][j], ords[p][m - k + j]
                if f[s] < f[t]:
                    ex[t], ef[t] = X[s], f[s]
        assert np.array_equal(_bits(X1), _bits(ex)) and np.array_equal(_bits(f1), _bits(ef)), trial
        assert np.all(f1 <= f), "a destination row got worse"


def _xfer_tasks(rs):
    spec = [(6, -5.0, 5.0), (4, -100.0, 100.0), (6, 0.0, 50.0)]
    return [dict(func=0, dim=D, pop=16, lo=lo, hi=hi, shift=rs.uniform(lo, hi, D).astype(f32)) for D, lo, hi in spec]


def check_transfer(orc):
    """A transfer event: donor elites placed on their box corners map exactly onto the
    receiver's corners (lo1 + (v - lo2) / r2 * r1), every other candidate gene is the
    receiver's own rank-j elite gene, and candidate j replaces own rank pop-k+j iff
    strictly better."""
    rs = np.random.default_rng(9)
    k, P = 3, 16
    tasks = _xfer_tasks(rs)
    for tcr in (0.0, 0.4):
        eng = orc.Engine(tasks, seed=9, transfer_interval=1, transfer_count=k, transfer_cr=tcr)
        pre = []
        for t, tk in enumerate(tasks):
            D, lo, hi = tk["dim"], tk["lo"], tk["hi"]
            X = rs.uniform(lo, hi, (P, D)).astype(f32)
            # the elite-to-be rows sit on box corners; they are the best by fitness
            X[:k] = np.where(rs.integers(0, 2, (k, D)) == 1, f32(hi), f32(lo))
            f = orc.eval32(tk, X)
            f[:k] = -1.0 - np.arange(k, dtype=f32)  # ranks 0..k-1 are rows k-1..0
            eng.set_population(t, X, f)
            pre.append((X, f))
        gen = 4
        eng.phase_transfer(gen)
        for t, tk in enumerate(tasks):
            X, f = pre[t]
            X1, f1 = eng.population(t)
            w = orc.draw_block(9, t, gen, P_MIG, 0, 0, 0)[0]
            dn = w % (len(tasks) - 1)
            dn += 1 if dn >= t else 0
            dX, df = pre[dn]
            own, don = _rank(f), _rank(df)
            lo1, hi1, lo2, hi2 = tk["lo"], tk["hi"], tasks[dn]["lo"], tasks[dn]["hi"]
            ex, ef = X.copy(), f.copy()
            for j in range(k):
                c = X[own[j]].copy()
                for d in range(min(tk["dim"], tasks[dn]["dim"])):
                    wu = orc.draw_block(9, t, gen, P_XFER, j, d & 1, (d >> 1) >> 2)[(d >> 1) & 3]
                    if tcr > 0 and f32(((wu >> 8) + 1) * 2.0 ** -24) <= f32(tcr):
                        continue
                    v = dX[don[j], d]
                    c[d] = f32(lo1) if v == f32(lo2) else f32(hi1)  # corners onto corners
                    assert v in (f32(lo2), f32(hi2))
                fc = orc.eval32(tk, c)[0]
                tgt = own[P - k + j]
                if fc < f[tgt]:
                    ex[tgt], ef[tgt] = c, fc
            assert np.array_equal(_bits(X1), _bits(ex)) and np.array_equal(_bits(f1), _bits(ef)), (tcr, t)


def check_cvrp_migration(orc):
    """CVRP: island q's rank-j tour replaces island q+1's rank m-k+j tour iff strictly cheaper."""
    rs = np.random.default_rng(13)
    n_c, n, m, k = 10, 4, 8, 2
    xy = rs.integers(0, 101, (n_c + 1, 2)).astype(np.int32)
    dem = np.concatenate([[0], rs.integers(1, 11, n_c)]).astype(np.int32)
    ins = dict(n=n_c, capacity=30, pop=n * m, xy=xy, demand=dem)
    eng = orc.CvrpEngine([ins], n_islands=n, migrate_interval=1, migrate_count=k, seed=3)
    for trial in range(8):
        tours = np.array([rs.permutation(n_c) + 1 for _ in range(n * m)], np.uint8)
        cost = rs.integers(100, 106, n * m).astype(np.int32)
        eng.set_population(0, tours, cost)
        eng.phase_migrate(0)
        t1, c1 = eng.population(0)
        et, ec = tours.copy(), cost.copy()
        ords = [[q * m + int(r) for r in np.lexsort((np.arange(m), cost[q * m:(q + 1) * m]))] for q in range(n)]
        for q in range(n):
            p = (q + 1) % n
            for j in range(k):
                s, t = ords[q][j], ords[p][m - k + j]
                if cost[s] < cost[t]:
                    et[t], ec[t] = tours[s], cost[s]
        assert np.array_equal(t1, et) and np.array_equal(c1, ec), trial


def test_exchange_is_a_rotation_per_slot(oracle_mod):
    check_exchange(oracle_mod)


def test_migration_rank_pairs_and_strictness(oracle_mod):
    check_migration(oracle_mod)


def test_transfer_box_map_and_targets(oracle_mod):
    check_transfer(oracle_mod)


def test_cvrp_migration_rank_pairs(oracle_mod):
    check_cvrp_migration(oracle_mod)


def test_exchange_keeps_the_best_over_a_run(oracle_mod):
    """A run with an exchange every generation never loses a task's best row."""
    from emtf import tasks as tg
    tasks = tg.benchmark_tasks(4, 10, 32, rotate=True, seed=2)
    eng = oracle_mod.Engine(tasks, seed=4, n_islands=4, exchange_interval=1)
    best = [eng.population(t)[1].min() for t in range(4)]
    for _ in range(5):
        eng.run(1)
        for t in range(4):
            b = eng.population(t)[1].min()
            assert b <= best[t]
            best[t] = b


def check_online_finished(o):
    test_online.test_finished_tasks_are_not_donors(o)


def check_online_new(o):
    test_online.test_new_and_pending_tasks_are_not_donors_transfer_crosses_batches(o)


def check_online_pick(o):
    test_online.test_donor_is_the_kth_other_donor(o)


def check_online_late(o):
    test_online.test_late_task_runs_its_own_generations(o)


# ------------------------------------------------------------- planted mistakes
MUTANTS = {
    1: ("exchange offset drawn per (island, slot)", check_exchange),
    2: ("migration writes rank m-1-j", check_migration),
    3: ("migration on 'better or equal'", check_migration),
    4: ("migration sends the source's worst", check_migration),
    5: ("transfer box map with the ranges swapped", check_transfer),
    6: ("transfer replaces own rank j", check_transfer),
    7: ("CVRP migration writes rank m-1-j", check_cvrp_migration),
    8: ("online: a task done at the boundary is still a donor", check_online_finished),
    9: ("online: a task admitted at the boundary is already a donor", check_online_new),
    10: ("online: the donor skips the task's id, not its list position", check_online_pick),
    11: ("online: draws addressed by the clock, not the task's generation", check_online_late),
}


@pytest.fixture(scope="module")
def mutant_libs(oracle_mod, tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    with ThreadPoolExecutor(4) as ex:
        paths = list(ex.map(lambda k: oracle_mod.build_mutant(k, str(d / f"liboracle_m{k}.so")), sorted(MUTANTS)))
    return dict(zip(sorted(MUTANTS), paths))


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_planted_mistake_is_caught(oracle_mod, mutant_libs, k):
    what, check = MUTANTS[k]
    with oracle_mod.swapped(mutant_libs[k]):
        with pytest.raises(AssertionError):
            check(oracle_mod)
