"""Array plumbing between the f2py-style call surface and the device.

torch tensors are the carrier: a call given torch device tensors runs
asynchronously on the current stream and returns device tensors; a call
given numpy arrays copies them to the current device, runs, waits, and
returns numpy arrays -- the drop-in behaviour of the reference's f2py
modules (float32/int8 casts, logical C-order indexing, zero-filled outputs).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("ct2 needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def is_host(*arrays):
    """True when no argument is a torch tensor (host round trip mode)."""
    return not any(isinstance(a, torch.Tensor) for a in arrays)


def to_dev(x, dtype, shape=None):
    """Contiguous device tensor of dtype (torch dtype) with the given shape."""
    if isinstance(x, torch.Tensor):
        t = x.to(device=device(), dtype=dtype)
    else:
        np_dt = {torch.float32: np.float32, torch.float16: np.float16, torch.int8: np.int8,
                 torch.float64: np.float64, torch.int32: np.int32}[dtype]
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np_dt)).to(device())
    t = t.contiguous()
    if shape is not None:
        if t.numel() != int(np.prod(shape)):
            raise ValueError(f"array of {t.numel()} elements does not match shape {tuple(shape)}")
        t = t.reshape(shape)
    return t


def ind_dev(ind, shape):
    """Indicator as int8 on the device, or None (every ray used)."""
    if ind is None:
        return None
    if isinstance(ind, torch.Tensor) and ind.dtype == torch.int8 and ind.is_cuda:
        return ind.contiguous().reshape(shape)
    # the kernels test `== 1` on int8; map any other integer type through the
    # same comparison first so e.g. 257 (== 1 after an int8 wrap) stays unused
    if isinstance(ind, torch.Tensor):
        return to_dev((ind == 1).to(torch.int8), torch.int8, shape)
    a = np.asarray(ind)
    if a.dtype != np.int8:
        a = (a == 1).astype(np.int8)
    return to_dev(a, torch.int8, shape)


def ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def out(t, host):
    if host:
        return t.cpu().numpy()
    return t
