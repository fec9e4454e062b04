"""The engine's random draws, exported from the device (emtf_eval_draws).

`eval_draws(seed, addrs)` returns, for every address row {task, generation,
purpose, index, sub, block} of an int64/int32 tensor addrs[n, 6], the four
32-bit words of the Philox4x32-10 block the engine draws there, computed by the
device's own generator on torch's current stream.
"""
import ctypes

from . import _lib

# purposes of the engine's draws (the counter's top byte)
P_INIT, P_DE_IDX, P_DE_CR, P_GA_MATE, P_GA_SBX, P_GA_PM, P_MIG, P_XFER, P_ISLX = range(1, 10)


def _bind():
    L = _lib.lib()
    if not getattr(L, "_draws_bound", False):
        vp = ctypes.c_void_p
        L.emtf_eval_draws.argtypes = [ctypes.c_uint64, ctypes.c_int32, vp, vp, vp]
        L.emtf_eval_draws.restype = ctypes.c_int
        L._draws_bound = True
    return L


def eval_draws(seed, addrs, device="cuda"):
    """uint32 words [n, 4] (as an int64 CPU tensor) of the blocks at addrs[n, 6]."""
    import torch
    a = torch.as_tensor(addrs, dtype=torch.int64)
    if a.dim() != 2 or a.shape[1] != 6:
        raise ValueError("addrs must have shape [n, 6]")
    if a.numel() and (a.min() < 0 or a.max() > 0xFFFFFFFF):
        raise ValueError("addresses are 32-bit unsigned")
    n = a.shape[0]
    # uint32 bit patterns carried in int32 storage
    dev = (a - (a >= 2 ** 31).to(torch.int64) * 2 ** 32).to(torch.int32).contiguous().to(device)
    out = torch.empty((max(n, 1), 4), dtype=torch.int32, device=device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(out.device).cuda_stream)
    _lib.check(_bind().emtf_eval_draws(ctypes.c_uint64(seed), n, ctypes.c_void_p(dev.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), stream))
    torch.cuda.synchronize(out.device)
    return out[:n].cpu().to(torch.int64) & 0xFFFFFFFF
