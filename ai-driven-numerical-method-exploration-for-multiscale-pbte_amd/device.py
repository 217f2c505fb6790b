"""Device buffers: PyTorch-ROCm tensors used only as HBM allocations and a
stream handle for the native library."""
from __future__ import annotations

import numpy as np

__all__ = ["require_gpu", "torch", "dev_zeros", "to_dev", "ptr", "stream_handle", "DEVICE"]

try:
    import torch
except ImportError as exc:  # pragma: no cover
    raise RuntimeError("PyTorch-ROCm is required for device buffers") from exc

DEVICE = "cuda:0"


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the solver runs only on the GPU "
                           "(there is no CPU fallback)")
    return torch.device(DEVICE)


def dev_zeros(shape, dtype=torch.float64):
    return torch.zeros(shape, dtype=dtype, device=require_gpu())


def to_dev(a: np.ndarray, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(require_gpu())


def ptr(t) -> int:
    if t is None:
        return 0
    assert t.is_cuda and t.is_contiguous()
    return t.data_ptr()


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream
