"""Device and stream plumbing (PyTorch-ROCm owns memory and streams)."""
from __future__ import annotations

import torch

from . import _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("fem_taichi_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    _lib.lib()


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    """hipStream_t of torch's current stream, as an int for ctypes."""
    return torch.cuda.current_stream().cuda_stream


def sync():
    torch.cuda.current_stream().synchronize()
