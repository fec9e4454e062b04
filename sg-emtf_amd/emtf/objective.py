"""Objectives and their hand-written adjoints as torch operators (ROCm).

`objective(func, x, rot=None, shift=None)` evaluates f(Q (x - o)) for a batch
x[n, dim] of float32 CUDA tensors through emtf_objective_fwd and
differentiates through emtf_objective_adjoint (include/emtf.h).  The kernels
run on torch's current stream.
"""
import ctypes

from . import _lib

f32p = ctypes.POINTER(ctypes.c_float)


def _bind():
    L = _lib.lib()
    if not getattr(L, "_obj_bound", False):
        vp = ctypes.c_void_p
        L.emtf_objective_fwd.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp, vp]
        L.emtf_objective_adjoint.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp, vp, vp]
        L._obj_bound = True
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _check_input(x):
    import torch
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32 and x.dim() == 2):
        raise TypeError("x must be a float32 CUDA tensor of shape [n, dim]")


def _operand(t, x, shape, name):
    """rot [dim, dim] / shift [dim] as a contiguous float32 tensor on x's device: the kernels
    read them as row-major binary32, so any other layout, dtype or device is refused."""
    import torch
    if t is None:
        return None
    if not (isinstance(t, torch.Tensor) and t.dtype == torch.float32 and t.device == x.device):
        raise TypeError(f"{name} must be a float32 tensor on {x.device}")
    if tuple(t.shape) != shape:
        raise ValueError(f"{name} must have shape {list(shape)}, got {list(t.shape)}")
    return t.contiguous()


def forward_raw(func, x, rot=None, shift=None):
    import torch
    _check_input(x)
    x = x.contiguous()
    D = x.shape[1]
    rot, shift = _operand(rot, x, (D, D), "rot"), _operand(shift, x, (D,), "shift")
    f = torch.empty(x.shape[0], device=x.device, dtype=torch.float32)
    stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(_bind().emtf_objective_fwd(int(func), x.shape[1], x.shape[0], _ptr(x), _ptr(rot), _ptr(shift),
                                          _ptr(f), stream))
    return f


def adjoint_raw(func, x, gf, rot=None, shift=None):
    import torch
    _check_input(x)
    x = x.contiguous()
    D = x.shape[1]
    rot, shift = _operand(rot, x, (D, D), "rot"), _operand(shift, x, (D,), "shift")
    gf = gf.contiguous().to(torch.float32)
    gx = torch.empty_like(x)
    stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(_bind().emtf_objective_adjoint(int(func), x.shape[1], x.shape[0], _ptr(x), _ptr(rot), _ptr(shift),
                                              _ptr(gf), _ptr(gx), stream))
    return gx


def _make_function():
    import torch

    class Objective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, func, rot, shift):
            ctx.func = func
            ctx.save_for_backward(x, rot, shift)
            return forward_raw(func, x, rot, shift)

        @staticmethod
        def backward(ctx, gf):
            x, rot, shift = ctx.saved_tensors
            return adjoint_raw(ctx.func, x, gf, rot, shift), None, None, None

    return Objective


_Objective = None


def objective(func, x, rot=None, shift=None):
    """Differentiable f(Q (x - o)) per row of x (float32 CUDA)."""
    global _Objective
    if _Objective is None:
        _Objective = _make_function()
    return _Objective.apply(x, int(func), rot, shift)
