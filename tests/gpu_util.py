"""Device-buffer helpers for the GPU parity tests (torch supplies the memory;
the kernels are called through the C ABI of libalien_hip.so).

`call` keeps every tensor argument referenced until the device has finished
the launch (it synchronises before returning), so a buffer created inline in
the argument list cannot be freed and reused while a kernel still reads it.
"""
import ctypes as C

import numpy as np
import torch

import alien_amd as aa


def dev(a: np.ndarray) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    torch.cuda.synchronize()
    return t


def empty(n, dtype=torch.float64) -> torch.Tensor:
    fill = float("nan") if dtype.is_floating_point else (0x5A if dtype == torch.uint8 else -7)
    t = torch.full((max(int(n), 1),), fill, dtype=dtype, device="cuda")
    torch.cuda.synchronize()
    return t


def host(t: torch.Tensor, ctx) -> np.ndarray:
    ctx.synchronize()
    torch.cuda.synchronize()
    return t.cpu().numpy()


def p(t: torch.Tensor) -> torch.Tensor:
    """Marks a device-buffer argument of `call` (converted there, kept alive there)."""
    assert isinstance(t, torch.Tensor) and t.is_cuda
    return t


def call(name, *args):
    conv = [C.c_void_p(a.data_ptr()) if isinstance(a, torch.Tensor) else a for a in args]
    rc = getattr(aa.lib(), name)(*conv)
    torch.cuda.synchronize()  # waits for every stream, ours included, while `args` holds the tensors
    aa.check(rc)


def row_scale(ro, cols, vals, x):
    """sum_j |a_ij x_j| per row: the scale of the per-row 1e-13 tolerance."""
    n = len(ro) - 1
    rows = np.repeat(np.arange(n), np.diff(ro))
    s = np.zeros(n)
    np.add.at(s, rows, np.abs(vals * x[cols]))
    return s
