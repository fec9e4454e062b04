"""The persistent schedule's ordering protocol, modelled on the CPU (tests/persist_model.c):
host threads stand in for workgroups, fewer of them than items, with random stalls.  The
protocol must finish with every mailbox read seeing the donor's previous epoch, and each
of its three waits must be necessary: taking one out (a planted mistake) must be caught."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("persist") / "persist_model")
    subprocess.run(["gcc", "-O2", "-pthread", "-o", out, os.path.join(HERE, "persist_model.c")], check=True)
    return out


def _run(model, *args):
    r = subprocess.run([model] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout.strip(), r.stderr


@pytest.mark.parametrize("tasks,epochs,threads", [(16, 20, 4), (5, 30, 5), (8, 30, 1), (32, 10, 12), (3, 40, 8)])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_protocol_orders_every_read(model, tasks, epochs, threads, seed):
    rc, out, err = _run(model, tasks, epochs, threads, seed)
    assert rc == 0 and out == "OK", err


@pytest.mark.parametrize("mistake", [1, 2, 3])
def test_each_wait_is_needed(model, mistake):
    """1: no donor wait; 2: publishing without waiting for the previous epoch's readers;
    3: no wait for the task's own previous epoch."""
    caught = [_run(model, 8, 30, 6, seed, mistake)[1] == "FAIL" for seed in range(1, 7)]
    assert any(caught)
