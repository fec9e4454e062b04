"""Gradients of the objectives.

CPU: the oracle's FP64 analytic gradients against FP64 central differences
(h = 1e-6) and against an independent FP64 torch-autograd implementation of the
textbook objectives (mpmath differences for Weierstrass, whose 3^20 frequency
defeats h = 1e-6).  GPU: emtf_objective_fwd bit-exact with the oracle, and the
binary32 adjoint (emtf_objective_adjoint, through torch.autograd) against the
FP64 gradients.
"""
import math

import numpy as np
import pytest
import torch

E = math.e


def torch_objective(func, z):
    """Textbook objectives in torch (FP64 autograd reference); z: [n, D]."""
    D = z.shape[1]
    j = torch.arange(D, dtype=z.dtype)
    if func == 0:
        return (z * z).sum(1)
    if func == 1:
        return (100 * (z[:, :-1] ** 2 - z[:, 1:]) ** 2 + (z[:, :-1] - 1) ** 2).sum(1)
    if func == 2:
        return (-20 * torch.exp(-0.2 * torch.sqrt((z * z).sum(1) / D)) - torch.exp(torch.cos(2 * math.pi * z).sum(1) / D)
                + 20 + E)
    if func == 3:
        return (z * z - 10 * torch.cos(2 * math.pi * z) + 10).sum(1)
    if func == 4:
        return 1 + (z * z).sum(1) / 4000 - torch.prod(torch.cos(z / torch.sqrt(j + 1)), 1)
    if func == 5:
        k = torch.arange(21, dtype=z.dtype)
        w0 = (0.5 ** k * torch.cos(math.pi * 3 ** k)).sum()
        return (0.5 ** k * torch.cos(2 * math.pi * 3 ** k * (z[..., None] + 0.5))).sum((1, 2)) - D * w0
    if func == 6:
        return 418.9828872724338 * D - (z * torch.sin(torch.sqrt(torch.abs(z)))).sum(1)
    if func == 7:
        c = 1e6 ** (j / (D - 1)) if D > 1 else torch.ones(1, dtype=z.dtype)
        return (c * z * z).sum(1)
    raise ValueError(func)


def _task(func, D, rotate, seed):
    rng = np.random.default_rng(seed)
    t = dict(func=func, dim=D, xform=1 if rotate else 0, shift=rng.uniform(-80, 80, D).astype(np.float32))
    if rotate:
        t["rot"] = np.linalg.qr(rng.standard_normal((D, D)))[0].astype(np.float32)
    return t


def _points(t, n, seed, scale):
    rng = np.random.default_rng(seed)
    return (t["shift"][None, :] + rng.normal(0, scale, (n, t["dim"]))).astype(np.float32)


@pytest.mark.parametrize("func", [0, 1, 2, 3, 4, 6, 7])
def test_grad64_vs_central_differences(oracle_mod, func):
    t = _task(func, 12, True, func)
    x = _points(t, 8, 1, 3.0).astype(np.float64)
    g = oracle_mod.grad64(t, x)
    h = 1e-6
    for d in range(12):
        e = np.zeros(12)
        e[d] = h
        fd = (oracle_mod.eval64(t, x + e) - oracle_mod.eval64(t, x - e)) / (2 * h)
        assert np.allclose(fd, g[:, d], rtol=1e-5, atol=1e-5 * (1 + np.abs(g).max())), (func, d)


@pytest.mark.parametrize("func", range(8))
def test_grad64_vs_torch_autograd(oracle_mod, func):
    t = _task(func, 10, True, 10 + func)
    x = _points(t, 6, 2, 0.3 if func == 5 else 3.0).astype(np.float64)
    X = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    Q = torch.tensor(t["rot"], dtype=torch.float64)
    o = torch.tensor(t["shift"], dtype=torch.float64)
    f = torch_objective(func, (X - o) @ Q.T)
    f.sum().backward()
    g = oracle_mod.grad64(t, x)
    assert np.allclose(X.grad.numpy(), g, rtol=1e-9, atol=1e-9 * (1 + np.abs(g).max()))
    assert np.allclose(f.detach().numpy(), oracle_mod.eval64(t, x), rtol=1e-10, atol=1e-9)


def test_weierstrass_grad64_vs_mpmath(oracle_mod):
    import mpmath as mp
    mp.mp.dps = 50
    t = _task(5, 3, False, 3)
    x = _points(t, 2, 3, 0.01).astype(np.float64)
    g = oracle_mod.grad64(t, x)
    for i in range(2):
        for d in range(3):
            def F(v):
                z = [mp.mpf(float(x[i, k])) - mp.mpf(float(t["shift"][k])) for k in range(3)]
                z[d] = v - mp.mpf(float(t["shift"][d]))
                return mp.fsum(mp.mpf(0.5) ** k * mp.cos(2 * mp.pi * 3 ** k * (zz + mp.mpf(0.5)))
                               for zz in z for k in range(21))
            num = mp.diff(F, mp.mpf(float(x[i, d])))
            # float64 evaluates cos(2 pi 3^20 (z + 1/2)) with ~2e-6 rad phase error; the k = 20 term scales it by 2 pi 1.5^20
            assert abs(float(num) - g[i, d]) <= 2e-5 * (1 + abs(float(num)))


@pytest.mark.parametrize("func", range(8))
def test_grad32_vs_grad64(oracle_mod, func):
    """The oracle's binary32 gradient (the device adjoint's restatement) tracks FP64."""
    for rotate in (False, True):
        exact_z = func == 5 and rotate
        t = _task(func, 30, rotate and not exact_z, 400 + func)
        x = _points(t, 16, func, 1e-3 if func == 5 else (0.5 if func in (3, 4) else 5.0))
        g32 = oracle_mod.grad32(t, x).astype(np.float64)
        g64 = oracle_mod.grad64(t, x.astype(np.float64))
        err = np.linalg.norm(g32 - g64, axis=1)
        scale = np.linalg.norm(g64, axis=1) + 1e-3
        assert np.all(err <= 1e-4 * scale), (func, rotate, float((err / scale).max()))


# ----------------------------------------------------------------------- GPU
def _cuda(a):
    return torch.tensor(a, dtype=torch.float32, device="cuda:0") if a is not None else None


@pytest.mark.gpu
@pytest.mark.parametrize("D", [7, 50, 100, 128])
def test_gpu_forward_bit_exact(oracle_mod, D):
    from emtf.objective import forward_raw
    for func in range(8):
        for rotate in (False, True):
            t = _task(func, D, rotate, 100 + func)
            x = _points(t, 70, func, 20.0)
            print("case: fwd D", D, "func", func, "rotate", rotate, flush=True)
            f = forward_raw(func, _cuda(x), _cuda(t.get("rot")), _cuda(t["shift"])).cpu().numpy()
            ref = oracle_mod.eval32(t, x)
            assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), (func, rotate)


@pytest.mark.gpu
@pytest.mark.parametrize("D", [129, 150, 200, 256])
def test_gpu_forward_wide_dims_bit_exact(oracle_mod, D):
    """Dimensions past the LDS evaluator (Q and points read from global memory)."""
    from emtf.objective import forward_raw
    for func in range(8):
        for rotate in (False, True):
            t = _task(func, D, rotate, 700 + func)
            x = _points(t, 150, func, 20.0)
            print("case: wide fwd D", D, "func", func, "rotate", rotate, flush=True)
            f = forward_raw(func, _cuda(x), _cuda(t.get("rot")), _cuda(t["shift"])).cpu().numpy()
            ref = oracle_mod.eval32(t, x)
            assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), (func, rotate)


@pytest.mark.gpu
@pytest.mark.parametrize("D", [7, 50, 200])
def test_gpu_weierstrass_term_ranges_bit_exact(oracle_mod, D):
    """The Weierstrass term at the edges of its branches (z = 0, the 2^-32 bound of the
    fixed-point path and the small-|z| branch below it, |z| near and past 2^23): zero
    shift, so z = x exactly; every point mixes the categories over its dimensions."""
    from emtf.objective import forward_raw
    lo = np.float32(2.0 ** -32)
    cats = np.array([0.0, lo, np.nextafter(lo, np.float32(0)), np.nextafter(lo, np.float32(1)), 1e-10, 3e-12,
                     1e-20, 0.3, 3.7, 0.5, 1.0, 1e6, 8388607.5, 2.0 ** 23, 1e8], np.float32)
    rng = np.random.default_rng(D)
    n = 96
    x = cats[rng.integers(0, len(cats), (n, D))] * rng.choice(np.float32([-1, 1]), (n, D))
    x[0] = 0.0  # f = 0 exactly
    x[1] = cats[4]  # every coordinate on the small-|z| branch
    t = dict(func=5, dim=D, xform=0, shift=np.zeros(D, np.float32))
    print("case: weierstrass ranges D", D, flush=True)
    f = forward_raw(5, _cuda(x), None, _cuda(t["shift"])).cpu().numpy()
    ref = oracle_mod.eval32(t, x)
    assert ref[0] == 0.0 and ref[1] > 0.0
    assert np.array_equal(f.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("D", [7, 50, 100, 128])
def test_gpu_adjoint_bit_exact(oracle_mod, D):
    """emtf_objective_adjoint with unit upstream gradient equals the oracle's grad32 bit for bit."""
    from emtf.objective import adjoint_raw
    for func in range(8):
        for rotate in (False, True):
            t = _task(func, D, rotate, 500 + func)
            x = _points(t, 70, func, 1e-3 if func == 5 else 3.0)
            print("case: adjoint bits D", D, "func", func, "rotate", rotate, flush=True)
            gf = torch.ones(70, dtype=torch.float32, device="cuda:0")
            g = adjoint_raw(func, _cuda(x), gf, _cuda(t.get("rot")), _cuda(t["shift"])).cpu().numpy()
            ref = oracle_mod.grad32(t, x)
            assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), \
                (func, rotate, int(np.sum(g.view(np.uint32) != ref.view(np.uint32))))


@pytest.mark.gpu
@pytest.mark.parametrize("D", [10, 50, 100])
def test_gpu_adjoint_vs_fp64(oracle_mod, D):
    from emtf.objective import objective
    for func in range(8):
        exact_z = func == 5  # Weierstrass: only where z = x - o is exact in binary32
        t = _task(func, D, not exact_z, 200 + func)
        x = _points(t, 64, func, 1e-3 if exact_z else (0.5 if func in (3, 4) else 5.0))
        print("case: adjoint D", D, "func", func, flush=True)
        X = _cuda(x).requires_grad_(True)
        f = objective(func, X, _cuda(t.get("rot")), _cuda(t["shift"]))
        gf = torch.linspace(0.5, 1.5, 64, device="cuda:0")
        (f * gf).sum().backward()
        g32 = X.grad.cpu().numpy().astype(np.float64)
        g64 = oracle_mod.grad64(t, x.astype(np.float64)) * gf.cpu().numpy()[:, None].astype(np.float64)
        err = np.linalg.norm(g32 - g64, axis=1)
        scale = np.linalg.norm(g64, axis=1) + 1e-3
        assert np.all(err <= 2e-3 * scale), (func, float((err / scale).max()))


@pytest.mark.gpu
def test_gpu_autograd_fd_check(oracle_mod):
    """FP64 central differences of the textbook objective against the autograd gradient."""
    from emtf.objective import objective
    D = 20
    for func in (0, 1, 2, 3, 7):
        t = _task(func, D, True, 300 + func)
        x = _points(t, 64, 7, 2.0)
        X = _cuda(x).requires_grad_(True)
        objective(func, X, _cuda(t["rot"]), _cuda(t["shift"])).sum().backward()
        g = X.grad.cpu().numpy().astype(np.float64)
        h = 1e-6
        xd = x.astype(np.float64)
        for d in range(0, D, 5):
            e = np.zeros(D)
            e[d] = h
            fd = (oracle_mod.eval64(t, xd + e) - oracle_mod.eval64(t, xd - e)) / (2 * h)
            assert np.allclose(g[:, d], fd, rtol=2e-3, atol=2e-3 * (1 + np.abs(fd).max())), (func, d)


def test_objective_operands_are_checked():
    """rot / shift reach the kernels as contiguous float32 on x's device, else are refused."""
    from emtf.objective import _operand
    x = torch.zeros(3, 4)
    rot = torch.arange(16, dtype=torch.float32).reshape(4, 4)
    r = _operand(rot.t(), x, (4, 4), "rot")
    assert r.is_contiguous() and torch.equal(r, rot.t())
    with pytest.raises(TypeError):
        _operand(rot.double(), x, (4, 4), "rot")
    with pytest.raises(ValueError):
        _operand(torch.zeros(3), x, (4,), "shift")
    assert _operand(None, x, (4,), "shift") is None


@pytest.mark.gpu
def test_gpu_transposed_rotation_view():
    """A transposed view of Q gives the same f as its contiguous copy."""
    from emtf.objective import forward_raw
    rs = np.random.default_rng(4)
    D = 30
    q = torch.tensor(np.linalg.qr(rs.standard_normal((D, D)))[0], dtype=torch.float32, device="cuda:0")
    x = torch.tensor(rs.uniform(-5, 5, (40, D)), dtype=torch.float32, device="cuda:0")
    a = forward_raw(3, x, q.t())
    b = forward_raw(3, x, q.t().contiguous())
    assert torch.equal(a, b)
