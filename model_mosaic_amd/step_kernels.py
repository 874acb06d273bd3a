"""Element-wise kernels of step3_propagation.py:94-110 on the HIP library."""
from __future__ import annotations

import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, check_out, require_gpu


def psi_from_proj(delta_raw, inv_norm, inv_beta_ratio, out=None) -> torch.Tensor:
    """psi = exp(-phase * inv_beta_ratio) * e^{i phase}, phase = delta_raw * inv_norm (float32 -> complex64)."""
    require_gpu()
    dev = delta_raw.device if isinstance(delta_raw, torch.Tensor) and delta_raw.is_cuda else torch.device(
        "cuda", torch.cuda.current_device())
    p = as_device_tensor(delta_raw, torch.float32, dev)
    if out is None:
        out = torch.empty(p.shape, dtype=torch.complex64, device=dev)
    else:
        check_out(out, p.shape, torch.complex64, dev, "psi_from_proj")
    _lib.check(_lib.load().mm_psi_from_proj(p.data_ptr(), out.data_ptr(), p.numel(), float(inv_norm),
                                            float(inv_beta_ratio), _lib.stream_handle()), "psi_from_proj")
    return out


def abs2(z, out=None) -> torch.Tensor:
    """|z|^2 of a complex64 array as float32."""
    require_gpu()
    dev = z.device if isinstance(z, torch.Tensor) and z.is_cuda else torch.device("cuda", torch.cuda.current_device())
    x = as_device_tensor(z, torch.complex64, dev)
    if out is None:
        out = torch.empty(x.shape, dtype=torch.float32, device=dev)
    else:
        check_out(out, x.shape, torch.float32, dev, "abs2")
    _lib.check(_lib.load().mm_abs2(x.data_ptr(), out.data_ptr(), x.numel(), _lib.stream_handle()), "abs2")
    return out
