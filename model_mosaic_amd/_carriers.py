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


def check_out(out, shape, dtype: torch.dtype, device: torch.device, what: str) -> torch.Tensor:
    """Validates a caller-supplied output buffer before its address reaches a kernel: a contiguous tensor
    of `dtype` on `device` with exactly prod(shape) elements (any shape of that size).  Raises ValueError
    otherwise; returns `out`."""
    if not isinstance(out, torch.Tensor):
        raise ValueError(f"{what}: out must be a torch tensor, got {type(out).__name__}")
    count = 1
    for d in shape:
        count *= int(d)
    if out.dtype != dtype:
        raise ValueError(f"{what}: out dtype {out.dtype} != {dtype}")
    if out.device != device:
        raise ValueError(f"{what}: out is on {out.device}, expected {device}")
    if not out.is_contiguous():
        raise ValueError(f"{what}: out must be contiguous")
    if out.numel() != count:
        raise ValueError(f"{what}: out has shape {tuple(out.shape)}, expected {tuple(int(d) for d in shape)}")
    return out


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("model_mosaic_amd operators need a ROCm GPU (torch.cuda.is_available() is False)")


def pinned_empty(shape, dtype: torch.dtype) -> torch.Tensor:
    """Page-locked host tensor (allocated pinned, never registered after the fact)."""
    return torch.empty(shape, dtype=dtype, pin_memory=True)
