"""Device plumbing: torch owns device buffers and streams; libmkgpu computes.

Helpers here turn tensors into raw pointers for the C ABI, hand out the
current HIP stream, and keep per-device scratch (the FP64 reduction
workspace and a small scalar block) so entry points never allocate on the
hot path.
"""
import threading

import numpy as np
import torch

from .lib import mkgpu

_tls = threading.local()


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("minkasi_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    """Current torch stream as a hipStream_t (void*)."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t, dtype=None, name="tensor"):
    """Raw device pointer of a contiguous CUDA/HIP tensor (checked)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch tensor on the GPU, got %s" % (name, type(t)))
    if not t.is_cuda:
        raise ValueError("%s must live on the GPU" % name)
    if dtype is not None and t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    return t.data_ptr()


def to_device(a, dtype=None):
    """numpy / tensor -> contiguous device tensor (no copy when already there)."""
    dev = device()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
        if dtype is not None:
            t = t.to(dtype)
        return t.contiguous()
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a).to(dev)
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


class Scratch:
    """Per-device reduction workspace + FP64 scalar block."""

    def __init__(self, dev):
        nbytes = int(mkgpu.lib.mk_reduce_workspace_bytes())
        self.work = torch.empty(nbytes // 8 + 32, dtype=torch.float64, device=dev)
        self.sc = torch.zeros(mkgpu.SC_COUNT, dtype=torch.float64, device=dev)


_scratch = {}


def scratch():
    dev = device()
    s = _scratch.get(dev.index)
    if s is None:
        s = Scratch(dev)
        _scratch[dev.index] = s
    return s


def torch_dtype(x):
    if isinstance(x, torch.Tensor):
        return x.dtype
    return {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}[np.asarray(x).dtype]
