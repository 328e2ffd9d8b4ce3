"""Device plumbing: the per-process lepic context and torch-owned HBM buffers.

torch provides device memory and the stream; every computation is a call into
liblepic_hip.so on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np
import torch

from . import _native

_ctx = {}


def require_gpu(device=None) -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("essaitemporaire_lepic_module_amd needs a ROCm GPU (torch.cuda is unavailable)")
    return torch.device(device if device is not None else "cuda:0")


def context(device=None) -> _native.Context:
    """The lepic context of `device` for the calling thread, bound to torch's
    current stream (which is per thread too).  A context holds the library's
    scratch and control buffers, so threads that drive the device at once
    (a sort on a side stream beside the field solve) each get their own."""
    dev = require_gpu(device)
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    key = (idx, threading.get_ident())
    ctx = _ctx.get(key)
    if ctx is None:
        with torch.cuda.device(idx):
            ctx = _native.Context(idx)
        _ctx[key] = ctx
    ctx.bind_stream(torch.cuda.current_stream(idx).cuda_stream)
    return ctx


def ptr(t) -> int:
    if t is None:
        return 0
    if not t.is_contiguous():
        raise ValueError("lepic needs contiguous tensors")
    return t.data_ptr()


def as_device(a, dtype, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


def dbl_array(vals, n=_native.MAX_SPECIES):
    arr = (C.c_double * n)()
    for i, v in enumerate(vals):
        arr[i] = float(v)
    return arr


def node_shape(n):
    return (int(n[2]) + 3, int(n[1]) + 3, int(n[0]) + 3)
