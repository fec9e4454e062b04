"""Benchmark: task-generations/s of the island engine on one MI355X.

Workload (the "plain GPU" benchmark configuration): 1024 tasks cycling
Sphere/Rastrigin/Ackley/Griewank/Rosenbrock/Schwefel/Weierstrass/Elliptic,
D = 50, z = Q (x - o) with Q from a seeded QR, o ~ U[-80, 80], box
[-100, 100], pop 128 per task, DE/rand/1/bin F = 0.5 CR = 0.9, 4 elites
migrating from a uniformly random other task every 20 generations, FP32.

A step is one epoch of 20 generations of every task (one kernel launch);
the default times 100 steps = the full 2000-generation run with every input
already resident in HBM.  The CPU baseline is the C oracle (OpenMP across
tasks) timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sg-emtf_amd"))

METRIC = "task-generations/s @1 GPU; time-to-1e-6 error, batched+adjoint vs plain DE"
PEAK_FP32_MFMA_TFLOPS = 157.3
# LDS: 128 bytes per clock per CU, 256 CUs at 2.4 GHz
PEAK_LDS_TBS = 256 * 128 * 2.4e9 / 1e12


def cpu_model():
    """The host CPU's model name (the CPU baseline's cores)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed epochs of --interval generations")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tasks", type=int, default=1024)
    ap.add_argument("--dim", type=int, default=50)
    ap.add_argument("--pop", type=int, default=128)
    ap.add_argument("--interval", type=int, default=20, help="generations per epoch = migration interval")
    ap.add_argument("--elites", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-gens", type=int, default=None,
                    help="generations of the CPU-baseline sample (0 = skip; default: plain 2000 (GA 400), uneven "
                         "1000 (GA 200), cvrp 2000 -- 10 to 20 s of CPU work on 16 threads)")
    ap.add_argument("--cpu-tasks", type=int, default=1024)
    ap.add_argument("--workload", default="plain",
                    choices=["plain", "uneven", "cvrp", "polish", "online", "convergence"],
                    help="plain: the rotated continuous workload (default); uneven: 4096 tasks with "
                         "D~U{10..200}, pop~U{32..256}; cvrp: the CVRP GA workload; "
                         "polish: time to mean error 1e-6, adjoint polish vs plain DE at D=100; "
                         "online: 10000 plain tasks admitted in 100 batches at U[0,100) ms gaps; "
                         "convergence: per-objective FP64 error statistics of the plain configuration "
                         "(--tasks/--dim/--pop/--gens) over seeds 0..N-1")
    ap.add_argument("--batches", type=int, default=100, help="online workload: arrival batches")
    ap.add_argument("--algo", default="de", choices=["de", "ga"],
                    help="plain / uneven workloads: DE/rand/1/bin (default) or the reference's GA operators "
                         "(SBX eta_c 2 + polynomial mutation eta_m 5, pm 1)")
    ap.add_argument("--seeds", type=int, default=10, help="polish / convergence workloads: seeds 0..N-1")
    ap.add_argument("--no-rotate", action="store_true",
                    help="convergence workload: unrotated tasks (the CPU configuration: --tasks 8 --pop 64 "
                         "--gens 1000 --interval 10 --elites 2 --no-rotate)")
    ap.add_argument("--no-tune", action="store_true", help="switch the kernel-variant / stream autotuner off")
    ap.add_argument("--gens", type=int, default=None,
                    help="polish / convergence workloads: generations per run (default 3000 / 2000)")
    ap.add_argument("--only-func", default=None, help="give every task this objective (per-objective timing)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reference for HBM counters)")
    ap.add_argument("--no-study", action="store_true",
                    help="plain workload: skip the one-seed time-to-1e-6 study (adjoint polish vs plain DE)")
    ap.add_argument("--one-call", action="store_true",
                    help="run the timed steps as one emtf_run call (the persistent schedule's single launch)")
    ap.add_argument("--schedule", default="auto", choices=["auto", "epochs", "persistent"],
                    help="auto: epochs replayed from a captured hipGraph; epochs: launched directly; "
                         "persistent: one launch per run, epochs ordered on the device")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain torch device copy of known size, profiled beside the kernel to scale FETCH/WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4, "launches": 5}))


def cpu_baseline(tasks, params, gens, n_tasks):
    """Oracle (OpenMP across tasks) on the first n_tasks tasks for `gens` generations."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    eng = oracle.Engine(tasks[:n_tasks], **{k: v for k, v in params.items() if k not in ("tune", "schedule")})
    t0 = time.perf_counter()
    eng.run(gens, threads)
    dt = time.perf_counter() - t0
    eng.close()
    return {"value": n_tasks * gens / dt, "unit": "task-generations/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n_tasks} tasks x {gens} generations of the same workload (oracle, C/OpenMP, "
                      f"{threads} threads, {dt:.2f} s)"}


def bench_cvrp(a):
    """CVRP GA workload: 512 instances, n ~ U{50..200} customers on integer U[0,100]^2, depot (50,50),
    demand U{1..10}, capacity 50, pop 128, 4 islands with 2 elites every 25 generations; a step is
    100 generations (one launch), 50 steps = 5000 generations."""
    from emtf import hip
    from emtf.cvrp import CvrpEngine, random_instances
    n_inst = 512
    steps = a.steps if a.steps != 100 else 50
    inst = random_instances(n_inst, seed=a.seed)
    params = dict(p_swap=0.3, p_2opt=0.3, n_islands=4, migrate_interval=25, migrate_count=2, seed=a.seed)
    eng = CvrpEngine(inst, **params)
    for _ in range(a.warmup):
        eng.run(100)
    hip.device_synchronize()
    evs = [hip.Event() for _ in range(steps + 1)]
    evs[0].record()
    for k in range(steps):
        eng.run(100)
        evs[k + 1].record()
    hip.device_synchronize()
    ms = sum(evs[k].elapsed_ms(evs[k + 1]) for k in range(steps))
    value = n_inst * 100 * steps / (ms / 1e3)
    best, _ = eng.best()
    cpu = None
    if a.cpu_gens > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        o = oracle.CvrpEngine(inst, **params)
        g = a.cpu_gens
        t0 = time.perf_counter()
        o.run(g, threads)
        dt = time.perf_counter() - t0
        cpu = {"value": n_inst * g / dt, "unit": "instance-generations/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"{n_inst} instances x {g} generations (oracle, C/OpenMP, {threads} threads, {dt:.2f} s)"}
    # algorithmic LDS bytes per offspring (two parents and the child of n one-byte genes,
    # the split's gathers over routes of about 9 customers, the mutation): about 12 n
    lds_bytes = sum(12 * ins["n"] * ins["pop"] for ins in inst) * 100
    achieved = lds_bytes / (ms / steps / 1e3) / 1e12
    out = {"metric": "CVRP instance-generations/s @1 GPU", "value": value, "unit": "instance-generations/s",
           "n_gpus": 1, "steps": steps, "warmup": a.warmup, "ms_per_step": ms / steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
           "config": {"workload": "CVRP GA: 512 instances, n~U{50..200}, capacity 50, pop 128, 4 islands, "
                                  "2 elites every 25 gens", "generations": 100 * steps, "generations_per_step": 100},
           "roofline": {"bound": "lds", "achieved": achieved, "peak": PEAK_LDS_TBS, "unit": "TB/s",
                        "frac": achieved / PEAK_LDS_TBS, "traffic": None,
                        "algorithmic": "12 n bytes of LDS per offspring (parents + child + split gathers), "
                                       "summed over instances and 100 generations per launch",
                        "note": "the split is a dependent chain per individual (latency-bound), far below "
                                "the LDS bandwidth"},
           "cpu_baseline": cpu, "mean_best_cost": float(best.mean())}
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out))


def _converge(tasks, params, gens, chunk, target):
    """Run in chunks of `chunk` generations; GPU time (HIP events) and FP64 error after each."""
    import numpy as np
    import emtf
    from emtf import hip
    from emtf.convergence import error64
    eng = emtf.Engine(tasks, **params)
    hip.device_synchronize()
    t_ms, errs = [0.0], [error64(tasks, eng.best()[1])]
    for _ in range(gens // chunk):
        e0, e1 = hip.Event(), hip.Event()
        e0.record()
        eng.run(chunk)
        e1.record()
        hip.device_synchronize()
        t_ms.append(t_ms[-1] + e0.elapsed_ms(e1))
        errs.append(error64(tasks, eng.best()[1]))
    eng.close()
    errs = np.array(errs)  # [points, tasks]
    return np.array(t_ms), errs


def _time_to(t_ms, mean_err, target):
    import numpy as np
    hit = np.nonzero(mean_err <= target)[0]
    return float(t_ms[hit[0]]) if len(hit) else None


NAMES = ["sphere", "rosenbrock", "ackley", "rastrigin", "griewank", "weierstrass", "schwefel", "elliptic"]


def time_to_target_study(seed=0, gens=3000, chunk=50, target=1e-6):
    """The second metric on one seed: the adjoint-polish configuration (1024 rotated tasks,
    D = 100, pop 128, DE/rand/1/bin, 4 elites every 20 gens; polish: 5 Armijo steps on the
    top 4 elites every 50 gens) run plain and polished.  Error = float64 f(best) - f* of
    every task (Schwefel excluded: its minimum lies outside the box).  Per objective: GPU
    time (kernels, HIP events) until the mean error <= target, for both variants, and the
    time ratio polished / plain where both reach it."""
    import numpy as np
    import emtf
    from emtf import tasks as tg
    tasks = tg.benchmark_tasks(1024, 100, 128, rotate=True, seed=seed)
    funcs = np.array([t["func"] for t in tasks])
    base = dict(algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=20, transfer_count=4, seed=seed)
    runs = {}
    for mode in ("plain", "polish"):
        p = dict(base)
        if mode == "polish":
            p.update(polish_interval=50, polish_steps=5, polish_topk=4)
        runs[mode] = _converge(tasks, p, gens, chunk, target)
    per = {}
    ratios = []
    for f, n in enumerate(NAMES):
        if n == "schwefel":
            continue
        row = {}
        for mode, (t_ms, errs) in runs.items():
            m = np.mean(errs[:, funcs == f], axis=1)
            row[mode + "_ms"] = _time_to(t_ms, m, target)
            row[mode + "_final_mean_error"] = float(m[-1])
        if row["plain_ms"] and row["polish_ms"]:
            row["ratio"] = row["polish_ms"] / row["plain_ms"]
            ratios.append(row["ratio"])
        per[n] = row
    live = funcs != NAMES.index("schwefel")
    full = {m: float(r[0][-1]) for m, r in runs.items()}
    mean_all = {m: _time_to(r[0], np.nanmean(r[1][:, live], axis=1), target) for m, r in runs.items()}
    return {"config": "1024 rotated tasks, D=100, pop 128, 3000 gens, seed %d; polish 5 Armijo steps x top 4 "
                      "every 50 gens" % seed,
            "target_mean_error": target, "per_objective": per,
            "ratio_geomean_over_objectives_reached_by_both": float(np.exp(np.mean(np.log(ratios)))) if ratios else None,
            "objectives_reached_by_both": [n for n, r in per.items() if "ratio" in r],
            "mean_error_all_objectives_ms": mean_all,
            "gpu_ms_full_run": full, "polish_overhead_frac": full["polish"] / full["plain"] - 1.0}


def bench_polish(a):
    """Adjoint-polish workload: 1024 rotated tasks, D=100, pop 128, DE/rand/1/bin with 4 elites
    transferred every 20 generations; the polished variant adds 5 Armijo steps on the top 4
    elites of every task every 50 generations.  Error = float64 re-evaluation of each task's
    best point (Schwefel excluded: no known minimum inside the box).  Reports the GPU time to
    mean error <= 1e-6 of both variants over seeds 0..N-1 and the per-objective picture
    (--only-func: every task of that objective, for per-objective kernel timing)."""
    import numpy as np
    import emtf
    from emtf import tasks as tg
    target, chunk = 1e-6, 50
    gens = a.gens if a.gens is not None else 3000
    names = ["sphere", "rosenbrock", "ackley", "rastrigin", "griewank", "weierstrass", "schwefel", "elliptic"]
    base = dict(algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=20, transfer_count=4)
    res = {"plain": [], "polish": []}
    per_func = {m: {n: [] for n in names} for m in res}
    final = {m: [] for m in res}
    for seed in range(a.seeds):
        only = [emtf.FUNC_IDS[a.only_func]] * 1024 if a.only_func else None
        tasks = tg.benchmark_tasks(1024, 100, 128, rotate=True, seed=seed, funcs=only)
        funcs = np.array([t["func"] for t in tasks])
        for mode in res:
            p = dict(base, seed=seed)
            if mode == "polish":
                p.update(polish_interval=50, polish_steps=5, polish_topk=4)
            t_ms, errs = _converge(tasks, p, gens, chunk, target)
            live = funcs != 6
            mean_err = np.nanmean(errs[:, live], axis=1) if live.any() else np.full(errs.shape[0], np.nan)
            res[mode].append(_time_to(t_ms, mean_err, target))
            final[mode].append((float(mean_err[-1]), float(t_ms[-1])))
            for f, n in enumerate(names):
                if f == 6:
                    continue
                m = np.mean(errs[:, funcs == f], axis=1)
                per_func[mode][n].append((_time_to(t_ms, m, target), float(np.mean(errs[-1, funcs == f] <= target)),
                                          float(m[-1])))
            print(f"seed {seed} {mode}: mean error {mean_err[-1]:.3e} after {gens} gens, "
                  f"{t_ms[-1]:.0f} ms GPU, time to 1e-6: {res[mode][-1]}", file=sys.stderr, flush=True)

    def summary(mode):
        tt = [v for v in res[mode] if v is not None]
        return {"runs_reaching": len(tt), "mean_ms_to_target": float(np.mean(tt)) if tt else None,
                "final_mean_error": float(np.mean([e for e, _ in final[mode]])),
                "gpu_ms_full_run": float(np.mean([t for _, t in final[mode]]))}
    funcs_out = {}
    for n in names:
        if n == "schwefel":
            continue
        row = {}
        for mode in res:
            tt = [t for t, _, _ in per_func[mode][n] if t is not None]
            row[mode] = {"runs_reaching": len(tt), "mean_ms_to_target": float(np.mean(tt)) if tt else None,
                         "frac_tasks_at_target": float(np.mean([fr for _, fr, _ in per_func[mode][n]])),
                         "final_mean_error": float(np.mean([e for _, _, e in per_func[mode][n]]))}
        funcs_out[n] = row
    # objectives whose mean error reaches the target in every run of both variants
    both = [n for n, r in funcs_out.items()
            if r["plain"]["runs_reaching"] == a.seeds and r["polish"]["runs_reaching"] == a.seeds]
    subset_ratio = (float(np.exp(np.mean([np.log(funcs_out[n]["polish"]["mean_ms_to_target"] /
                                                 funcs_out[n]["plain"]["mean_ms_to_target"]) for n in both])))
                    if both else None)
    sp, sq = summary("plain"), summary("polish")
    ratio = (sq["mean_ms_to_target"] / sp["mean_ms_to_target"]
             if sp["mean_ms_to_target"] and sq["mean_ms_to_target"] else None)
    out = {"metric": "GPU time to mean error 1e-6: adjoint polish / plain DE", "value": ratio, "unit": "ratio",
           "n_gpus": 1, "steps": gens // chunk, "warmup": 0, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "adjoint polish: 1024 rotated tasks, D=100, pop 128, DE/rand/1/bin F=0.5 CR=0.9, "
                                  "4 elites every 20 gens; polish 5 Armijo steps on the top 4 elites every 50 gens",
                      "generations": gens, "seeds": a.seeds, "target": target,
                      "error": "float64 f(best) - f*, mean over non-Schwefel tasks"},
           "plain": sp, "polish": sq, "per_objective": funcs_out,
           "ratio_on_objectives_reached_by_both": {"objectives": both, "geomean_time_ratio": subset_ratio}}
    print(json.dumps(out))


def bench_convergence(a):
    """Convergence statistics of the plain configuration over seeds 0..N-1 (default: 1024
    rotated tasks, D=50, pop 128, DE/rand/1/bin F=0.5 CR=0.9, 4 elites every 20 gens, 2000
    generations; --tasks 8 --pop 64 --gens 1000 --interval 10 --elites 2 --no-rotate gives the
    CPU configuration: 8 unrotated tasks, 2 elites every 10 gens).  Error =
    float64 f(best) - f* of every task after every 100 generations (Schwefel excluded: its
    minimum lies outside the box).  Per objective, over the seeds: the final mean error
    (mean and standard deviation), the fraction of tasks at 1e-6 and the GPU time until the
    objective's mean error reaches 1e-6."""
    import numpy as np
    import emtf
    from emtf import tasks as tg
    gens = a.gens if a.gens is not None else 2000
    chunk, target = 100, 1e-6
    rows = {n: [] for n in NAMES if n != "schwefel"}
    gpu_ms = []
    for seed in range(a.seeds):
        tasks = tg.benchmark_tasks(a.tasks, a.dim, a.pop, rotate=not a.no_rotate, seed=seed)
        funcs = np.array([t["func"] for t in tasks])
        p = dict(algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=a.interval, transfer_count=a.elites,
                 seed=seed)
        t_ms, errs = _converge(tasks, p, gens, chunk, target)
        gpu_ms.append(float(t_ms[-1]))
        for f, n in enumerate(NAMES):
            if n == "schwefel" or not np.any(funcs == f):
                continue
            m = np.mean(errs[:, funcs == f], axis=1)
            rows[n].append((float(m[-1]), float(np.mean(errs[-1, funcs == f] <= target)), _time_to(t_ms, m, target)))
        print(f"seed {seed}: {gens} generations, {t_ms[-1]:.1f} ms GPU", file=sys.stderr, flush=True)
    per = {}
    for n, r in rows.items():
        if not r:
            continue
        fin = np.array([x[0] for x in r])
        tt = [x[2] for x in r if x[2] is not None]
        per[n] = {"final_mean_error_mean": float(fin.mean()), "final_mean_error_std": float(fin.std()),
                  "frac_tasks_at_1e-6": float(np.mean([x[1] for x in r])), "runs_reaching_1e-6": len(tt),
                  "mean_ms_to_1e-6": float(np.mean(tt)) if tt else None}
    out = {"metric": "per-objective FP64 error after the run, mean over seeds", "value": None, "unit": "error",
           "n_gpus": 1, "steps": gens // chunk, "warmup": 0, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"plain: {a.tasks} {'unrotated' if a.no_rotate else 'rotated'} tasks, D={a.dim}, pop={a.pop}, DE/rand/1/bin F=0.5 CR=0.9, "
                                  f"{a.elites} elites every {a.interval} gens",
                      "generations": gens, "seeds": a.seeds, "target": target},
           "gpu_ms_per_run": float(np.mean(gpu_ms)), "per_objective": per}
    print(json.dumps(out))


def bench_online(a):
    """Online admission (the reference's driver, SG-EMTF.cu:1842-1958): T tasks of the plain
    configuration split into batches at random cut points, batch i arriving U[0, 100) ms after
    batch i-1; all tasks live in one engine with online admission, a batch is admitted at the
    clock the queued work has reached when it arrives and runs 2000 generations, drawing
    transfer donors among every running task.  Reports the mean response time (arrival ->
    admitted) and solve time (arrival -> finished) per task."""
    import numpy as np
    import emtf
    from emtf import online
    from emtf import tasks as tg
    T = a.tasks if a.tasks != 1024 else 10000
    gens = a.interval * a.steps
    tasks = tg.benchmark_tasks(T, a.dim, a.pop, rotate=True, seed=a.seed)
    cuts, arr = online.arrival_schedule(T, a.batches, max_gap_ms=100.0, seed=a.seed)
    params = dict(algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=a.interval, transfer_count=a.elites)
    # one untimed run loads the code objects
    warm = online.OnlineRun(tasks[:8], [0, 8], [0.0], a.interval, seed=a.seed, chunk=a.interval, **params)
    warm.run()
    warm.close()
    run = online.OnlineRun(tasks, cuts, arr, gens, seed=a.seed, chunk=a.interval, **params)
    resp, solve = run.run()
    best = run.best()
    run.close()
    span = run.makespan_ms
    out = {"metric": "online admission: mean solve time per task (arrival -> finished)", "value": float(solve.mean()),
           "unit": "ms", "n_gpus": 1, "steps": a.steps, "warmup": 1, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"online: {T} rotated tasks D={a.dim} pop={a.pop} in {a.batches} batches, "
                                  f"gaps U[0,100) ms, {gens} generations each, DE/rand/1/bin F=0.5 CR=0.9, "
                                  f"{a.elites} elites every {a.interval} gens from any running task",
                      "tasks": T, "batches": a.batches, "generations": gens},
           "mean_response_ms": float(resp.mean()), "max_response_ms": float(resp.max()),
           "mean_solve_ms": float(solve.mean()), "max_solve_ms": float(solve.max()),
           "last_arrival_ms": float(arr[-1]), "makespan_ms": span,
           "task_generations_per_s_over_makespan": T * gens / (span / 1e3),
           "mean_best_fitness": float(np.mean(best))}
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the engine shards tasks across "
              f"GPUs only by design, see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    if a.cpu_gens is None:
        ga = a.algo == "ga"
        a.cpu_gens = {"plain": 400 if ga else 2000, "uneven": 200 if ga else 1000, "cvrp": 2000}.get(a.workload, 100)
    if a.calibrate_copy:
        calibrate_copy(a.calibrate_copy)
        return
    if a.workload == "cvrp":
        bench_cvrp(a)
        return
    if a.workload == "polish":
        bench_polish(a)
        return
    if a.workload == "online":
        bench_online(a)
        return
    if a.workload == "convergence":
        bench_convergence(a)
        return
    import emtf
    from emtf import hip
    from emtf import tasks as tg

    uneven = a.workload == "uneven"
    if uneven:
        a.tasks = a.tasks if a.tasks != 1024 else 4096
        tasks = tg.uneven_tasks(a.tasks, seed=a.seed)
    else:
        funcs = [emtf.FUNC_IDS[a.only_func]] * a.tasks if a.only_func else None
        tasks = tg.benchmark_tasks(a.tasks, a.dim, a.pop, rotate=True, seed=a.seed, funcs=funcs)
    sched = {"auto": emtf.SCHEDULE_AUTO, "epochs": emtf.SCHEDULE_EPOCHS,
             "persistent": emtf.SCHEDULE_PERSISTENT}[a.schedule]
    if a.algo == "ga":
        params = dict(algo=emtf.GA_SBX_PM, eta_c=2.0, eta_m=5.0, pm=1.0, transfer_interval=a.interval,
                      transfer_count=a.elites, seed=a.seed, tune=1 if a.no_tune else 0, schedule=sched)
    else:
        params = dict(algo=emtf.DE_RAND_1_BIN, F=0.5, CR=0.9, transfer_interval=a.interval,
                      transfer_count=a.elites, seed=a.seed, tune=1 if a.no_tune else 0, schedule=sched)
    algo_name = "GA SBX(eta_c 2)+PM(eta_m 5, pm 1)" if a.algo == "ga" else "DE/rand/1/bin F=0.5 CR=0.9"
    eng = emtf.Engine(tasks, **params)
    for _ in range(a.warmup):
        eng.run(a.interval)
    hip.device_synchronize()

    evs = [hip.Event() for _ in range(a.steps + 1)]
    hip.device_synchronize()
    t0 = time.perf_counter()
    evs[0].record()
    if a.one_call:
        eng.run(a.interval * a.steps)
        evs[a.steps].record()
    else:
        for k in range(a.steps):
            eng.run(a.interval)  # one kernel launch on the null stream
            evs[k + 1].record()
    hip.device_synchronize()
    wall = time.perf_counter() - t0
    if a.one_call:
        launch_ms = [evs[0].elapsed_ms(evs[a.steps]) / a.steps] * a.steps
    else:
        launch_ms = [evs[k].elapsed_ms(evs[k + 1]) for k in range(a.steps)]
    gpu_ms = sum(launch_ms)
    task_gens = a.tasks * a.interval * a.steps
    value = task_gens / (gpu_ms / 1e3)

    flop_per_launch = 2.0 * sum(t["pop"] * t["dim"] ** 2 for t in tasks) * a.interval
    avg_launch_s = (gpu_ms / a.steps) / 1e3
    achieved = flop_per_launch / avg_launch_s / 1e12
    best, _ = eng.best()
    eng.close()

    # HBM bytes per step from the committed counter runs (FETCH_SIZE / WRITE_SIZE scaled by
    # a plain copy) when this is the configuration they were taken on
    traffic = None
    tp = os.path.join(ROOT, "profiles", "island_epoch_traffic.json")
    tu = os.path.join(ROOT, "profiles", "uneven_epoch_traffic.json")
    if (not uneven and not a.only_func and a.algo == "de" and (a.tasks, a.dim, a.pop, a.interval) == (1024, 50, 128, 20)
            and os.path.exists(tp)):
        traffic = json.load(open(tp))["hbm_bytes_per_launch"]
    if uneven and a.algo == "de" and (a.tasks, a.interval) == (4096, 20) and os.path.exists(tu):
        traffic = json.load(open(tu))["hbm_bytes_per_epoch"]

    cpu = None
    if a.cpu_gens > 0:
        cpu = cpu_baseline(tasks, params, a.cpu_gens, min(a.cpu_tasks, a.tasks))
    study = None
    if not uneven and not a.no_study and not a.only_func:
        study = time_to_target_study(seed=a.seed)

    out = {
        "metric": METRIC, "value": value, "unit": "task-generations/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": gpu_ms / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"batched uneven: {a.tasks} rotated tasks, D~U{{10..200}}, pop~U{{32..256}} rounded to 32"
                                if uneven else
                                f"plain GPU: {a.tasks} rotated tasks{' (all ' + a.only_func + ')' if a.only_func else ''}, D={a.dim}, pop={a.pop}")
                               + f", {algo_name}, {a.elites} elites every {a.interval} gens, "
                               f"{a.interval * a.steps} timed generations",
                   "tasks": a.tasks, "dim": "U{10..200}" if uneven else a.dim, "pop": "U{32..256}" if uneven else a.pop,
                   "generations": a.interval * a.steps,
                   "generations_per_step": a.interval, "parallelism": "tasks (1 GPU)"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_MFMA_TFLOPS, "traffic": traffic,
                     "algorithmic": "2*pop*D^2 FLOP per task-generation (rotation GEMM)"
                                    + (", summed over tasks" if uneven else ""),
                     "avg_launch_ms": gpu_ms / a.steps},
        "cpu_baseline": cpu,
        "time_to_1e-6": study,
        "launch": {"auto": "epochs replayed from a captured hipGraph", "epochs": "epochs launched directly",
                   "persistent": "persistent schedule, epochs ordered on the device"}[a.schedule]
                  + (f"; the {a.steps} steps as one run call" if a.one_call else ""),
        "wall_s": wall,
        "mean_best_fitness": float(best.mean()),
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
