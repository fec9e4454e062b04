"""emtf_get_elites_noalloc (the elite export staged in the library's fixed device array)
against emtf_get_elites, through the C ABI: bit-equal output on two calls in a row, the
second on a smaller input, and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind():
    from emtf import _lib
    L = _lib.lib()
    f32p = ctypes.POINTER(ctypes.c_float)
    for name in ("emtf_get_elites", "emtf_get_elites_noalloc"):
        fn = getattr(L, name)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int32, f32p, f32p, ctypes.c_int32]
        fn.restype = ctypes.c_int
    return L


def _export(L, name, eng, k, stride, with_x=True):
    f32p = ctypes.POINTER(ctypes.c_float)
    fit = np.full((eng.n_tasks, k), np.nan, np.float32)
    x = np.full((eng.n_tasks, k, stride), np.nan, np.float32) if with_x else None
    rc = getattr(L, name)(eng._h, k, fit.ctypes.data_as(f32p), x.ctypes.data_as(f32p) if with_x else None, stride)
    return rc, fit, x


def test_noalloc_elites_equal_the_allocating_export():
    import emtf
    from emtf import tasks as tg
    L = _bind()
    print("case: no-alloc elite export, 64 tasks D=50 then 16 tasks D=20", flush=True)
    big = emtf.Engine(tg.benchmark_tasks(64, 50, 128, rotate=True, seed=1), seed=1, transfer_interval=5,
                      transfer_count=2)
    big.run(10)
    small = emtf.Engine(tg.benchmark_tasks(16, 20, 64, rotate=False, seed=2), seed=2)
    small.run(7)
    for eng, k, stride in ((big, 8, 60), (small, 3, 20), (small, 5, 32)):
        rc_a, fa, xa = _export(L, "emtf_get_elites", eng, k, stride)
        rc_b, fb, xb = _export(L, "emtf_get_elites_noalloc", eng, k, stride)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, L.emtf_last_error())
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), (k, stride)
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), (k, stride)
        assert np.all(xb[:, :, eng.max_dim:] == 0) if stride > eng.max_dim else True
    # fitness only
    rc, fo, _ = _export(L, "emtf_get_elites_noalloc", small, 4, 20, with_x=False)
    rc2, fr, _ = _export(L, "emtf_get_elites", small, 4, 20, with_x=False)
    assert rc == 0 and rc2 == 0 and np.array_equal(fo.view(np.uint32), fr.view(np.uint32))
    # beyond the fixed array: 64 tasks x 128 elites x (255 + 1) floats > 2^20
    rc, _, _ = _export(L, "emtf_get_elites_noalloc", big, 128, 255)
    assert rc == -1 and "fixed array" in L.emtf_last_error().decode()
    big.close()
    small.close()
