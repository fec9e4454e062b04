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
    orc = oracle_mod.Engine(tasks, **params)
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
