"""C ABI checks that need no GPU: exported symbols and argument validation.

emtf_create validates every descriptor before touching the device, so the
error paths run on a host without a GPU.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|void)\s+\**\s*(emtf_\w+)\s*\(", src, re.M))
    return names


def test_library_exports_every_declared_symbol():
    from emtf import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert {"emtf_create", "emtf_run", "emtf_destroy", "emtf_get_population"} <= names
    for n in sorted(names):
        assert hasattr(lib, n), n


def _desc(func=0, dim=10, pop=32, xform=0, lo=-1.0, hi=1.0, rot=None, diag=None):
    from emtf import _lib
    f32p = _lib.f32p
    shift = np.zeros(dim, np.float32)
    keep = [shift, rot, diag]
    d = _lib.TaskDesc(func, dim, pop, xform, lo, hi,
                      rot.ctypes.data_as(f32p) if rot is not None else None, shift.ctypes.data_as(f32p),
                      diag.ctypes.data_as(f32p) if diag is not None else None)
    return d, keep


def _create(descs, **kw):
    from emtf import _lib
    from emtf.engine import DEFAULT_PARAMS
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    arr = (_lib.TaskDesc * len(descs))(*descs)
    h = ctypes.c_void_p()
    rc = _lib.lib().emtf_create(arr, len(descs), ctypes.byref(_lib.Params(**p)), ctypes.byref(h))
    return rc, _lib.lib().emtf_last_error().decode()


@pytest.mark.parametrize("kw, expect, msg", [
    (dict(func=9), -1, "unknown func"),
    (dict(dim=0), -1, "dim"),
    (dict(dim=257), -4, "dim 257"),
    (dict(pop=33), -1, "multiple of n_islands"),
    (dict(lo=1.0, hi=1.0), -1, "bad box"),
    (dict(xform=1), -1, "rotation matrix missing"),
    (dict(xform=2), -1, "diagonal missing"),
    (dict(pop=4096), -4, "pop 4096"),
])
def test_invalid_task_descriptors(kw, expect, msg):
    extra = {}
    if "pop" in kw and kw["pop"] == 33:
        extra = dict(n_islands=2)
    d, keep = _desc(**kw)
    rc, err = _create([d], **extra)
    assert rc == expect and msg in err, (rc, err)


@pytest.mark.parametrize("where", ["rot", "diag", "shift"])
@pytest.mark.parametrize("bad", [float("nan"), float("inf"), -float("inf")])
def test_non_finite_task_arrays_rejected(where, bad):
    """A NaN or infinity in the rotation, the diagonal or the shift is refused at create."""
    from emtf import _lib
    D = 6
    rot = np.eye(D, dtype=np.float32)
    diag = np.ones(D, np.float32)
    shift = np.zeros(D, np.float32)
    {"rot": rot, "diag": diag, "shift": shift}[where].ravel()[D + 1 if where == "rot" else 3] = bad
    f32p = _lib.f32p
    xform = {"rot": 1, "diag": 2, "shift": 1}[where]
    d = _lib.TaskDesc(0, D, 32, xform, -1.0, 1.0, rot.ctypes.data_as(f32p) if xform == 1 else None,
                      shift.ctypes.data_as(f32p), diag.ctypes.data_as(f32p) if xform == 2 else None)
    rc, err = _create([d])
    assert rc == -1 and "non-finite" in err, (rc, err)


def test_invalid_params():
    d, keep = _desc()
    assert _create([d], algo=7)[0] == -1
    assert "transfer_count" in _create([d, d], transfer_interval=5, transfer_count=20)[1]
    assert "migrate_count" in _create([d], n_islands=4, migrate_interval=1, migrate_count=5)[1]
    d2, keep2 = _desc(pop=30)
    assert "even island" in _create([d2], algo=2, n_islands=2)[1]


def test_online_admission_arguments():
    import emtf
    from emtf import tasks as tg
    d, keep = _desc()
    assert "admission" in _create([d], admission=3)[1]
    assert "online" in _create([d, d], admission=1, polish_interval=5, polish_steps=2, polish_topk=1)[1]
    from emtf import _lib
    L = _lib.lib()
    assert L.emtf_admit(None, 0, 1, 10, None) == -1
    del emtf, tg


def test_params_layout_matches_the_header(tmp_path):
    """The ctypes mirror of emtf_params has the C struct's size and field offsets."""
    import shutil
    import subprocess
    from emtf import _lib
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc missing")
    names = [f[0] for f in _lib.Params._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"emtf.h\"\nint main(void) {\n"
                   + "".join(f'    printf("%zu\\n", offsetof(emtf_params, {n}));\n' for n in names)
                   + '    printf("%zu\\n", sizeof(emtf_params));\n    return 0;\n}\n')
    exe = str(tmp_path / "layout")
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe], check=True)
    got = [int(v) for v in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(_lib.Params, n).offset for n in names] + [ctypes.sizeof(_lib.Params)]
    assert got == want


def test_null_arguments():
    from emtf import _lib
    L = _lib.lib()
    assert L.emtf_create(None, 1, None, None) == -1
    assert L.emtf_run(None, 5, None) == -1
    assert L.emtf_destroy(None) == 0


def test_persistent_schedule_limits():
    """The persistent schedule runs one LDS-kernel bucket without polish or admission; the
    other handles are refused at create, before any device work."""
    import emtf
    from emtf import tasks as tg
    tasks = tg.benchmark_tasks(2, 10, 32, rotate=False, seed=2)
    with pytest.raises(emtf.EmtfError, match="persistent schedule: gradient polish"):
        emtf.Engine(tasks, schedule=emtf.SCHEDULE_PERSISTENT, polish_interval=5, polish_steps=1, polish_topk=1)
    with pytest.raises(emtf.EmtfError, match="persistent schedule: online admission"):
        emtf.Engine(tasks, schedule=emtf.SCHEDULE_PERSISTENT, admission=emtf.ADMIT_ONLINE)
    two = tg.benchmark_tasks(2, 0, 32, rotate=False, seed=2, dims=[10, 100], pops=[32, 32])
    with pytest.raises(emtf.EmtfError, match="persistent schedule: every task must fall in one"):
        emtf.Engine(two, schedule=emtf.SCHEDULE_PERSISTENT, transfer_interval=5, transfer_count=2)
    wide = tg.benchmark_tasks(2, 150, 32, rotate=False, seed=2)
    with pytest.raises(emtf.EmtfError, match="persistent schedule: every task must fall in one"):
        emtf.Engine(wide, schedule=emtf.SCHEDULE_PERSISTENT)
    with pytest.raises(emtf.EmtfError, match="schedule"):
        emtf.Engine(tasks, schedule=7)


C_EXAMPLE = os.path.join(os.path.dirname(__file__), "c_abi", "abi_example.c")


def _build_c_example(tmp_path):
    import shutil
    import subprocess
    from emtf import _lib
    LIB = _lib.LIB_PATH
    gcc = shutil.which("gcc")
    if not gcc or not os.path.exists(LIB):
        pytest.skip("gcc or libemtf.so missing")
    exe = str(tmp_path / "abi_example")
    cmd = [gcc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), C_EXAMPLE, "-o", exe,
           "-L", os.path.dirname(LIB), "-lemtf", "-Wl,-rpath," + os.path.dirname(LIB)]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_c_example_compiles_against_header(tmp_path):
    """A plain C99 program builds against include/emtf.h and links libemtf.so."""
    _build_c_example(tmp_path)


@pytest.mark.gpu
def test_c_example_runs(tmp_path, oracle_mod):
    """The C caller's results equal the oracle's for the same tasks and parameters."""
    import subprocess
    import numpy as np
    exe = _build_c_example(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120).stdout.split()
    assert out[0] == "ok", out
    T, D = 16, 20
    tasks = []
    for t in range(T):
        sh = np.array([((t * 7 + d * 3) % 41) - 20.0 for d in range(D)], np.float32)
        tasks.append(dict(func=3 if t % 2 else 0, dim=D, pop=64, xform=0, lo=-100.0, hi=100.0, shift=sh))
    orc = oracle_mod.Engine(tasks, seed=7, F=0.5, CR=0.9, transfer_interval=20, transfer_count=2)
    orc.run(200)
    fo = [orc.population(t)[1] for t in range(2)]
    best = [np.float32(np.min(f)) for f in fo]
    # "%.9g" round-trips binary32
    assert [np.float32(v) for v in out[1:3]] == best
    assert int(out[3]) == 5 and np.float32(out[4]) == best[0]
    second = np.sort(fo[0])[1]  # the elite export's rank 1 of task 0
    assert np.float32(out[5]) == second


def test_trace_argument_checks():
    from emtf import _lib
    L = _lib.lib()
    n = ctypes.c_int32()
    assert L.emtf_trace_start(None, 5, 4, None) == -1
    assert L.emtf_trace(None, ctypes.byref(n), None, None) == -1


def test_eval_draws_argument_checks():
    from emtf import _lib
    from emtf.draws import _bind
    L = _bind()
    assert L.emtf_eval_draws(1, 0, None, None, None) == 0
    assert L.emtf_eval_draws(1, -1, None, None, None) == -1
    assert L.emtf_eval_draws(1, 4, None, None, None) == -1 and "NULL" in _lib.lib().emtf_last_error().decode()


@pytest.mark.gpu
def test_eval_draws_match_oracle(oracle_mod):
    """Device Philox blocks at random addresses equal the oracle's (pinned by rocRAND and
    Random123 known answers), including index / sub values past their 16 / 8 counter bits."""
    from emtf.draws import eval_draws
    rng = np.random.default_rng(5)
    n = 1000
    addrs = np.stack([rng.integers(0, 5000, n), rng.integers(0, 2 ** 32, n, dtype=np.int64), rng.integers(1, 18, n),
                      rng.integers(0, 70000, n), rng.integers(0, 300, n), rng.integers(0, 2 ** 32, n, dtype=np.int64)], 1)
    seed = 0xFEDCBA9876543210
    got = eval_draws(seed, addrs).numpy()
    for i in range(n):
        t, g, p, idx, sub, blk = (int(v) for v in addrs[i])
        want = oracle_mod.draw_block(seed, t, g, p, idx & 0xFFFF, sub & 0xFF, blk)
        assert list(got[i]) == want, (i, addrs[i])
