"""Bit-exact parity of the HIP engine (through the C ABI) with the CPU oracle.

Both sides consume the same Philox draws and follow the same binary32
arithmetic, so populations and fitness values must agree bit for bit after
any number of generations.
"""
import numpy as np
import pytest

import emtf
from emtf import tasks as tg

pytestmark = pytest.mark.gpu


def _compare(gpu, orc, tag):
    for t in range(gpu.n_tasks):
        xg, fg = gpu.population(t)
        xo, fo = orc.population(t)
        nbad = int(np.sum(fg.view(np.uint32) != fo.view(np.uint32)))
        assert nbad == 0, f"{tag}: task {t} fitness differs in {nbad} entries (first gpu {fg[:4]}, oracle {fo[:4]})"
        assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), f"{tag}: task {t} population differs"


def _pair(tasks, oracle_mod, **params):
    print(f"case: {len(tasks)} tasks dims {sorted(set(t['dim'] for t in tasks))} "
          f"pops {sorted(set(t['pop'] for t in tasks))} params {params}", flush=True)
    gpu = emtf.Engine(tasks, **params)
    orc = oracle_mod.Engine(tasks, **{k: v for k, v in params.items() if k not in ("tune", "schedule")})
    return gpu, orc


@pytest.mark.parametrize("dim", [50, 7, 32, 33, 64])
def test_init_bit_exact_all_objectives(oracle_mod, dim):
    tasks = tg.benchmark_tasks(8, dim, 64, rotate=True, seed=11)
    gpu, orc = _pair(tasks, oracle_mod, seed=5)
    _compare(gpu, orc, f"init D={dim}")


def test_de_rand_1_bin_rotated_with_transfer(oracle_mod):
    tasks = tg.benchmark_tasks(16, 50, 128, rotate=True, seed=3)
    gpu, orc = _pair(tasks, oracle_mod, seed=7, F=0.5, CR=0.9, transfer_interval=10, transfer_count=4)
    for step in (1, 9, 10, 15, 5):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"gen {orc.generation}")


def test_de_shift_only_pop64(oracle_mod):
    tasks = tg.benchmark_tasks(8, 50, 64, rotate=False, seed=4)
    gpu, orc = _pair(tasks, oracle_mod, seed=1, transfer_interval=10, transfer_count=2)
    gpu.run(25)
    orc.run(25)
    _compare(gpu, orc, "shift-only 25 gens")


def test_reference_profile_islands(oracle_mod):
    """DE/current/1, CR 0.6, diagonal transform, islands with exchange + migration + transfer."""
    tasks = tg.reference_tasks(6, 50, 128, seed=2)
    params = dict(seed=3, algo=emtf.DE_CURRENT_1_BIN, F=0.5, CR=0.6, n_islands=8, exchange_interval=1,
                  migrate_interval=5, migrate_count=4, transfer_interval=10, transfer_count=20, transfer_cr=0.6)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    gpu.run(23)
    orc.run(23)
    _compare(gpu, orc, "ref profile 23 gens")


def test_ga_sbx_pm_islands(oracle_mod):
    tasks = tg.reference_tasks(6, 50, 64, seed=8)
    params = dict(seed=9, algo=emtf.GA_SBX_PM, eta_c=2.0, eta_m=5.0, pm=1.0, n_islands=4, exchange_interval=1,
                  migrate_interval=5, migrate_count=2)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    gpu.run(12)
    orc.run(12)
    _compare(gpu, orc, "GA 12 gens")


@pytest.mark.parametrize("pop", [256, 96])
def test_multi_tile_and_ragged_population(oracle_mod, pop):
    tasks = tg.benchmark_tasks(8, 40, pop, rotate=True, seed=6)
    gpu, orc = _pair(tasks, oracle_mod, seed=2, transfer_interval=4, transfer_count=3)
    gpu.run(9)
    orc.run(9)
    _compare(gpu, orc, f"pop {pop}")


def test_uneven_dims_d100(oracle_mod):
    dims = [100, 17, 64, 65, 3, 128, 50, 99]
    tasks = tg.benchmark_tasks(8, 0, 64, rotate=True, seed=9, dims=dims)
    gpu, orc = _pair(tasks, oracle_mod, seed=4, transfer_interval=3, transfer_count=5)
    gpu.run(7)
    orc.run(7)
    _compare(gpu, orc, "uneven dims")


# ---- wide tasks (rows in global memory) and mixed kernel buckets
@pytest.mark.parametrize("dim", [129, 150, 200, 256])
def test_wide_dims_de_with_transfer(oracle_mod, dim):
    tasks = tg.benchmark_tasks(8, dim, 64, rotate=True, seed=40 + dim)
    gpu, orc = _pair(tasks, oracle_mod, seed=3, transfer_interval=5, transfer_count=3)
    for step in (1, 4, 6):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"wide D={dim} gen {orc.generation}")


def test_wide_from_lds_overflow(oracle_mod):
    """D = 100 with pop 1024 does not fit LDS and runs on the wide kernel."""
    tasks = tg.benchmark_tasks(4, 100, 1024, rotate=True, seed=9)
    gpu, orc = _pair(tasks, oracle_mod, seed=8, transfer_interval=3, transfer_count=8)
    gpu.run(4)
    orc.run(4)
    _compare(gpu, orc, "pop 1024 x D 100")


def test_wide_ga_islands(oracle_mod):
    tasks = tg.reference_tasks(4, 180, 64, seed=5)
    params = dict(seed=2, algo=emtf.GA_SBX_PM, eta_c=2.0, eta_m=5.0, pm=0.5, n_islands=4, exchange_interval=2,
                  migrate_interval=3, migrate_count=2)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    gpu.run(7)
    orc.run(7)
    _compare(gpu, orc, "wide GA 7 gens")


@pytest.mark.parametrize("algo", [emtf.DE_RAND_1_BIN, emtf.DE_CURRENT_1_BIN, emtf.GA_SBX_PM])
def test_uneven_batch_buckets(oracle_mod, algo):
    """D ~ U{10..200}, pop ~ U{32..256}: every kernel bucket at once, with islands and transfer."""
    tasks = tg.uneven_tasks(48, seed=4)
    params = dict(seed=6, algo=algo, n_islands=2, exchange_interval=4, migrate_interval=3, migrate_count=2,
                  transfer_interval=5, transfer_count=2, transfer_cr=0.3)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    for step in (5, 6):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"uneven algo {algo} gen {orc.generation}")


# ---- persistent schedule (one launch per run, device-side mailbox ordering; its ordering
# protocol is modelled on the CPU in test_persist_model.py)


@pytest.mark.parametrize("algo", [emtf.DE_RAND_1_BIN, emtf.GA_SBX_PM])
def test_persistent_schedule_bit_exact(oracle_mod, algo):
    tasks = tg.benchmark_tasks(24, 40, 64, rotate=True, seed=12)
    params = dict(seed=4, algo=algo, n_islands=2, exchange_interval=3, migrate_interval=4, migrate_count=2,
                  transfer_interval=5, transfer_count=3, transfer_cr=0.2)
    print(f"case: persistent schedule, 24 tasks D=40 pop 64, algo {algo}", flush=True)
    per = emtf.Engine(tasks, schedule=emtf.SCHEDULE_PERSISTENT, **params)
    epo = emtf.Engine(tasks, schedule=emtf.SCHEDULE_EPOCHS, **params)
    orc = oracle_mod.Engine(tasks, **params)
    for step in (1, 9, 10, 23):
        per.run(step)
        epo.run(step)
        orc.run(step)
        _compare(per, orc, f"persistent gen {orc.generation}")
        _compare(epo, orc, f"epochs gen {orc.generation}")


def test_persistent_many_epochs_per_call(oracle_mod):
    """More epochs than one persistent launch holds: the call is split into chunks."""
    tasks = tg.benchmark_tasks(6, 10, 32, rotate=False, seed=1)
    params = dict(seed=9, transfer_interval=1, transfer_count=2)
    print("case: persistent schedule, 6 tasks, 1030 one-generation epochs", flush=True)
    per = emtf.Engine(tasks, schedule=emtf.SCHEDULE_PERSISTENT, **params)
    orc = oracle_mod.Engine(tasks, **params)
    per.run(1030)
    orc.run(1030)
    _compare(per, orc, "1030 one-generation epochs")


def test_persistent_without_transfer_and_with_trace(oracle_mod):
    """No transfer (one epoch per launch), and launches cut at the convergence-trace points:
    the trace and the populations equal the epoch schedule's."""
    tasks = tg.benchmark_tasks(40, 20, 64, rotate=True, seed=5)
    print("case: persistent schedule without transfer / with trace", flush=True)
    for params in (dict(seed=3), dict(seed=3, transfer_interval=4, transfer_count=2)):
        per = emtf.Engine(tasks, schedule=emtf.SCHEDULE_PERSISTENT, **params)
        epo = emtf.Engine(tasks, schedule=emtf.SCHEDULE_EPOCHS, **params)
        per.trace_start(3, 10)
        epo.trace_start(3, 10)
        for step in (7, 11):
            per.run(step)
            epo.run(step)
        _compare(per, epo, f"persistent vs epochs {params}")
        assert np.array_equal(per.trace().view(np.uint32), epo.trace().view(np.uint32))


@pytest.mark.parametrize("tune", [0, 1, 2, 3])
def test_autotune_is_bit_exact(oracle_mod, tune):
    """The autotuner times the LDS kernel, the wide kernel and the wide kernel with Q in LDS
    of each bucket at create (and restores the state), then keeps one; every epoch must
    agree with the oracle whichever runs (tune 2 and 3 force the alternatives)."""
    tasks = tg.uneven_tasks(40, seed=8, dmax=220, pmax=256)
    params = dict(seed=3, transfer_interval=2, transfer_count=2, n_islands=2, migrate_interval=3, migrate_count=2)
    gpu = emtf.Engine(tasks, tune=tune, **params)
    orc = oracle_mod.Engine(tasks, **params)
    for step in (2, 2, 2, 2, 2, 6):
        gpu.run(step)
        gpu.sync()
        orc.run(step)
        _compare(gpu, orc, f"tune={tune} gen {orc.generation}")


_TUNE_CHILD = r"""
import os, sys
sys.path.insert(0, os.environ["EMTF_PKG"])
import numpy as np
import emtf
from emtf import tasks as tg
tasks = tg.uneven_tasks(24, seed=31, dmax=180, pmax=192)
eng = emtf.Engine(tasks, seed=4, tune=0, transfer_interval=3, transfer_count=2)
eng.run(6)
f = np.concatenate([eng.population(t)[1] for t in range(len(tasks))])
np.save(sys.argv[1], f)
"""


def test_tuner_decisions_are_cached(oracle_mod, tmp_path, monkeypatch, capfd):
    """The autotuner's decisions for a batch are reused: a second create in the same process
    and a create in a fresh process reading the EMTF_TUNE_CACHE file skip the calibration
    (the tuner's log says "cached"), pick the same variants and compute the oracle's bits."""
    import os
    import subprocess
    import sys
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, EMTF_TUNE_CACHE=str(cache), EMTF_TUNE_LOG="1",
               EMTF_PKG=os.path.dirname(os.path.dirname(emtf.__file__)))
    runs = []
    for k in range(2):
        print(f"case: tuner cache, fresh process {k}", flush=True)
        out = tmp_path / f"f{k}.npy"
        r = subprocess.run([sys.executable, "-c", _TUNE_CHILD, str(out)], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        runs.append((r.stderr, np.load(out)))
    lines0 = [l for l in runs[0][0].splitlines() if l.startswith("emtf tune:")]
    lines1 = [l for l in runs[1][0].splitlines() if l.startswith("emtf tune:")]
    assert lines0 and not any("(cached)" in l for l in lines0)
    assert lines1 and all(l.endswith("(cached)") for l in lines1)
    # the same decisions: identical lines but for the marker
    assert [l.replace(" (cached)", "") for l in lines1] == lines0
    assert len(cache.read_text().splitlines()) == 1
    tasks = tg.uneven_tasks(24, seed=31, dmax=180, pmax=192)
    orc = oracle_mod.Engine(tasks, seed=4, transfer_interval=3, transfer_count=2)
    orc.run(6)
    fo = np.concatenate([orc.population(t)[1] for t in range(len(tasks))])
    for _, f in runs:
        assert np.array_equal(f.view(np.uint32), fo.view(np.uint32))
    # in this process: the second create of a batch takes the first one's decisions
    monkeypatch.setenv("EMTF_TUNE_CACHE", "")
    monkeypatch.setenv("EMTF_TUNE_LOG", "1")
    tasks = tg.uneven_tasks(24, seed=32, dmax=180, pmax=192)
    orc = oracle_mod.Engine(tasks, seed=4, transfer_interval=3, transfer_count=2)
    orc.run(6)
    capfd.readouterr()
    a = emtf.Engine(tasks, seed=4, tune=0, transfer_interval=3, transfer_count=2)
    la = [l for l in capfd.readouterr().err.splitlines() if l.startswith("emtf tune:")]
    b = emtf.Engine(tasks, seed=4, tune=0, transfer_interval=3, transfer_count=2)
    lb = [l for l in capfd.readouterr().err.splitlines() if l.startswith("emtf tune:")]
    assert la and not any("(cached)" in l for l in la)
    assert lb and all(l.endswith("(cached)") for l in lb)
    a.run(6)
    b.run(6)
    _compare(a, orc, "tuned")
    _compare(b, orc, "tuned from the in-process cache")


@pytest.mark.parametrize("schedule", [emtf.SCHEDULE_AUTO, emtf.SCHEDULE_EPOCHS])
def test_graph_replay_bit_exact(oracle_mod, schedule):
    """Epochs replayed from the captured hipGraph (AUTO) and launched directly (EPOCHS) both
    equal the oracle: a batch of several launch units forked over streams, runs that start on
    and off transfer boundaries (replayed and direct epochs interleaved), a trace started in
    between, and host reads between replays."""
    dims = [20, 50, 50, 150, 100, 33, 64, 20] * 3
    pops = [32, 128, 256, 64, 96, 64, 160, 32] * 3
    tasks = tg.benchmark_tasks(len(dims), 0, 0, rotate=True, seed=21, dims=dims, pops=pops)
    gpu, orc = _pair(tasks, oracle_mod, seed=17, F=0.5, CR=0.9, transfer_interval=10, transfer_count=3,
                     schedule=schedule)
    for k, step in enumerate((10, 10, 5, 5, 20, 13, 7, 10)):
        if k == 5:
            gpu.trace_start(10, 4)
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"schedule {schedule} gen {orc.generation}")


def test_graph_replay_without_transfer(oracle_mod):
    """No transfer: every epoch is replayable; a new run length recaptures the graph."""
    tasks = tg.benchmark_tasks(12, 40, 64, rotate=True, seed=5)
    gpu, orc = _pair(tasks, oracle_mod, seed=3, F=0.5, CR=0.9)
    for step in (7, 7, 3, 7, 7):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"no transfer gen {orc.generation}")


def test_wide_512_threads_ga(oracle_mod):
    """The 512-thread wide build (tune=2 selects it for populations above four tiles)."""
    tasks = tg.benchmark_tasks(6, 170, 224, rotate=True, seed=13)
    params = dict(seed=5, algo=emtf.GA_SBX_PM, n_islands=4, exchange_interval=2, migrate_interval=3,
                  migrate_count=3, transfer_interval=3, transfer_count=4)
    gpu, orc = _pair(tasks, oracle_mod, tune=2, **params)
    gpu.run(7)
    orc.run(7)
    _compare(gpu, orc, "wide x512 GA")


@pytest.mark.parametrize("algo", [0, 2])
def test_wide_q_in_lds(oracle_mod, algo):
    """The wide kernel with the task's Q image staged in LDS (tune=3), NT 5 and 6 rows."""
    tasks = tg.benchmark_tasks(4, 170, 224, rotate=True, seed=17) + tg.benchmark_tasks(4, 150, 96, rotate=True, seed=18)
    params = dict(seed=6, algo=algo, n_islands=4, exchange_interval=2, migrate_interval=3,
                  migrate_count=3, transfer_interval=3, transfer_count=4)
    gpu, orc = _pair(tasks, oracle_mod, tune=3, **params)
    gpu.run(7)
    orc.run(7)
    _compare(gpu, orc, f"wide Q in LDS algo {algo}")


@pytest.mark.parametrize("algo", [emtf.DE_RAND_1_BIN, emtf.DE_CURRENT_1_BIN])
@pytest.mark.parametrize("m", [4, 5, 8])
def test_mate_rejection_small_islands(oracle_mod, algo, m):
    """Islands of m = 4, 5, 8 individuals: most mate draws collide, so the straight-line
    seven-word rejection and its sequential replay (more than seven words) both run, in
    the LDS and the wide kernel."""
    tasks = tg.benchmark_tasks(3, 40, 8 * m, rotate=True, seed=20 + m) + \
        tg.benchmark_tasks(2, 150, 8 * m, rotate=True, seed=30 + m)
    params = dict(seed=m, algo=algo, n_islands=8, exchange_interval=3, transfer_interval=4, transfer_count=2)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    for step in (5, 6):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"m={m} algo={algo} gen {orc.generation}")


def test_runs_on_changing_streams_stay_ordered(oracle_mod):
    """Consecutive runs queued on different streams without host synchronisation give the
    same populations as one stream (each run waits for the previous call's work)."""
    from emtf import hip
    tasks = tg.benchmark_tasks(12, 50, 128, rotate=True, seed=21)
    params = dict(seed=4, transfer_interval=5, transfer_count=2)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    streams = [hip.Stream() for _ in range(3)]
    for k in range(9):
        gpu.run(5, stream=streams[k % 3].handle)
    orc.run(45)
    _compare(gpu, orc, "45 generations over three streams")
    gpu.close()


@pytest.mark.parametrize("algo", [emtf.DE_RAND_1_BIN, emtf.DE_CURRENT_1_BIN, emtf.GA_SBX_PM])
def test_smallest_shapes(oracle_mod, algo):
    """D = 1, 2, 3 with populations of 2 to 8 (islands too small for three distinct mates
    take the deterministic fallback of the rejection draw), every objective, transfers."""
    tasks = (tg.benchmark_tasks(8, 1, 2, rotate=False, seed=31) + tg.benchmark_tasks(8, 2, 4, rotate=True, seed=32)
             + tg.benchmark_tasks(8, 3, 8, rotate=True, seed=33))
    params = dict(seed=7, algo=algo, transfer_interval=3, transfer_count=1)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    for step in (1, 4, 6):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"small shapes algo {algo} gen {orc.generation}")


def test_cpu_plain_configuration_full_run(oracle_mod):
    """The CPU-plain configuration in full: 8 unrotated tasks (the eight objectives), D = 50,
    pop 64, DE/rand/1/bin F = 0.5 CR = 0.9, 1000 generations, 2 elites from a random other
    task every 10 generations; bit-exact at generations 10, 500 and 1000."""
    tasks = tg.benchmark_tasks(8, 50, 64, rotate=False, seed=0)
    params = dict(seed=0, algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=10, transfer_count=2)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    for step in (10, 490, 500):
        gpu.run(step)
        orc.run(step)
        _compare(gpu, orc, f"CPU-plain configuration gen {orc.generation}")


def test_plain_gpu_configuration_full_run_subset(oracle_mod):
    """64 tasks of the plain GPU configuration (rotated, D = 50, pop 128, 4 elites every 20
    generations) over all 2000 generations, bit-exact at the end."""
    tasks = tg.benchmark_tasks(64, 50, 128, rotate=True, seed=0)
    params = dict(seed=0, algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=20, transfer_count=4)
    gpu, orc = _pair(tasks, oracle_mod, **params)
    gpu.run(2000)
    orc.run(2000, threads=8)
    _compare(gpu, orc, "plain GPU configuration, 64 tasks, gen 2000")


def _fitness_matches_fresh_evaluation(gpu, tasks, ids):
    """Stored fitness of every individual equals a fresh device evaluation of its row."""
    import torch
    from emtf.objective import forward_raw
    for t in ids:
        task = tasks[t]
        x, f = gpu.population(t)
        rot = torch.from_numpy(np.ascontiguousarray(task["rot"], np.float32)).cuda() if "rot" in task else None
        shift = torch.from_numpy(np.ascontiguousarray(task["shift"], np.float32)).cuda()
        fresh = forward_raw(task["func"], torch.from_numpy(x).cuda(), rot, shift).cpu().numpy()
        assert np.array_equal(fresh.view(np.uint32), f.view(np.uint32)), f"task {t}: stored fitness differs"


def test_full_size_plain_properties():
    """The plain GPU configuration at its real size (1024 tasks, 2000 generations): every
    task's best never increases (device trace every 100 generations) and every stored
    fitness equals a fresh evaluation of its row."""
    tasks = tg.benchmark_tasks(1024, 50, 128, rotate=True, seed=0)
    gpu = emtf.Engine(tasks, seed=0, algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=20, transfer_count=4)
    gpu.trace_start(100, 21)
    gpu.run(2000)
    rows = gpu.trace()
    assert rows.shape == (1024, 22)
    assert np.all(np.diff(rows[:, :-1], axis=1) <= 0)
    assert np.all(rows[:, -1] >= rows[:, 0])
    _fitness_matches_fresh_evaluation(gpu, tasks, range(0, 1024, 7))
    gpu.close()


def test_full_size_uneven_properties():
    """The batched uneven configuration at its real size (4096 tasks, D ~ U{10..200}, pop ~
    U{32..256}) over all 2000 generations: best never increases, stored fitness equals a
    fresh evaluation (every 31st task, the wide dimensions included)."""
    tasks = tg.uneven_tasks(4096, seed=0)
    gpu = emtf.Engine(tasks, seed=0, algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=20, transfer_count=4)
    gpu.trace_start(100, 21)
    gpu.run(2000)
    rows = gpu.trace()
    assert rows.shape == (4096, 22)
    assert np.all(np.diff(rows[:, :-1], axis=1) <= 0)
    ids = list(range(0, 4096, 31))
    assert sum(tasks[t]["dim"] > 128 for t in ids) > 30
    _fitness_matches_fresh_evaluation(gpu, tasks, ids)
    gpu.close()
