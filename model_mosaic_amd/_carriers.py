"""Conversion of caller arrays to the torch (ROCm) tensors the C ABI consumes."""
from __future__ import annotations

import numpy as np
import torch


def device_of(plan_device: int) -> torch.device:
    return torch.device("cuda", plan_device)


def as_device_tensor(x, dtype: torch.dtype, device: torch.device) -> torch.Tensor:
    """Contiguous tensor of `dtype` on `device` (copies only when needed)."""
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(x))
    if t.dtype != dtype:
        t = t.to(dtype)
    if t.device != device:
        t = t.to(device, non_blocking=False)
    return t.contiguous()


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("model_mosaic_amd operators need a ROCm GPU (torch.cuda.is_available() is False)")


def pinned_empty(shape, dtype: torch.dtype) -> torch.Tensor:
    """Page-locked host tensor (allocated pinned, never registered after the fact)."""
    return torch.empty(shape, dtype=dtype, pin_memory=True)
