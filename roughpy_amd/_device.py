"""Device plumbing: torch provides HBM allocations and the current HIP stream.

The compute path has no CPU fallback: without a ROCm GPU every compute call
raises RuntimeError.
"""
from __future__ import annotations

import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError(
            "roughpy_amd computes on a ROCm GPU (gfx950); no GPU is visible and "
            "there is no CPU fallback")
    return t


def is_device_array(a) -> bool:
    t = torch()
    return isinstance(a, t.Tensor) and a.is_cuda


def stream_handle():
    """hipStream_t of torch's current stream, as an integer for ctypes."""
    t = require_gpu()
    return t.cuda.current_stream().cuda_stream


def torch_dtype(np_dtype):
    t = torch()
    dt = np.dtype(np_dtype)
    if dt == np.float64:
        return t.float64
    if dt == np.float32:
        return t.float32
    raise TypeError(f"unsupported dtype {dt}")


def np_dtype(a) -> np.dtype:
    if isinstance(a, np.ndarray):
        return a.dtype
    t = torch()
    if isinstance(a, t.Tensor):
        return {t.float64: np.dtype(np.float64), t.float32: np.dtype(np.float32),
                t.int64: np.dtype(np.int64), t.int32: np.dtype(np.int32)}.get(
            a.dtype, np.dtype(object))
    return np.asarray(a).dtype


def to_device(a, dtype=None):
    """Contiguous device tensor holding `a` (no copy if already suitable)."""
    t = require_gpu()
    if isinstance(a, t.Tensor):
        out = a if a.is_cuda else a.cuda()
        if dtype is not None:
            out = out.to(torch_dtype(dtype))
        return out.contiguous()
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return t.from_numpy(arr).cuda()


def empty(shape, dtype):
    t = require_gpu()
    return t.empty(tuple(shape), dtype=torch_dtype(dtype), device="cuda")


def zeros(shape, dtype):
    t = require_gpu()
    return t.zeros(tuple(shape), dtype=torch_dtype(dtype), device="cuda")


def ptr(tensor) -> int:
    return tensor.data_ptr()


def to_host(tensor) -> np.ndarray:
    return tensor.detach().cpu().numpy()
