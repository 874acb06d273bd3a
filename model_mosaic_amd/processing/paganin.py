"""Paganin: single-distance Paganin phase retrieval on one MI355X.

Mirrors ``Paganin`` of processing/paganin.py:29-92: constructor
``Paganin(n, nz, ntheta, wavelength, voxelsize, distance, alpha)``,
``retrieve(intensity)`` with intensity (b <= ntheta, nz, n) or (nz, n)
float32, returning a new array of the same shape:

    H = alpha / (wavelength * distance * (kx^2 + ky^2) / (4 pi) + alpha)
    out = -log(max(irfft2(H * rfft2(I)), 1e-6))

All arithmetic happens in libmosaic_hip.so (mm_paganin_*): two real frames
ride in one complex 2-D transform (H is real and even), in three HBM passes
(row FFT, column FFT x H x column IFFT, row IFFT + clip + log).
"""
from __future__ import annotations

import ctypes as C

import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, check_out, require_gpu


class Paganin:
    """Single-distance Paganin phase retrieval sized for (ntheta x nz x n) intensities."""

    def __init__(self, n, nz, ntheta, wavelength, voxelsize, distance, alpha):
        require_gpu()
        self._lib = _lib.load()
        self.n = int(n)
        self.nz = int(nz)
        self._ntheta = int(ntheta)
        self.wavelength = float(wavelength)
        self.voxelsize = float(voxelsize)
        self.distance = float(distance)
        self.alpha = float(alpha)
        self.device = torch.device("cuda", torch.cuda.current_device())
        plan = C.c_void_p()
        _lib.check(self._lib.mm_paganin_create(C.byref(plan), self.n, self.nz, self._ntheta, self.wavelength,
                                               self.voxelsize, self.distance, self.alpha), "Paganin")
        self._plan = plan

    def free(self):
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            self._lib.mm_paganin_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def workspace_bytes(self) -> int:
        return int(self._lib.mm_paganin_workspace_bytes(self._plan)) if self._plan else 0

    def retrieve(self, intensity, out: torch.Tensor | None = None) -> torch.Tensor:
        """(b, nz, n) or (nz, n) float32 intensities -> phase of the same shape (processing/paganin.py:70)."""
        if not self._plan:
            raise RuntimeError("Paganin: plan was freed")
        x = as_device_tensor(intensity, torch.float32, self.device).contiguous()
        added_dim = x.ndim == 2
        if added_dim:
            x = x[None]
        if x.ndim != 3 or x.shape[1] != self.nz or x.shape[2] != self.n:
            raise ValueError(f"Paganin: intensity shape {tuple(x.shape)} != (b, {self.nz}, {self.n})")
        b = x.shape[0]
        assert b <= self._ntheta, f"batch {b} > cached ntheta {self._ntheta}"
        if out is None:
            res = out = torch.empty_like(x)
            if added_dim:
                res = out[0]
        else:
            res = check_out(out, x.shape, torch.float32, self.device, "Paganin.retrieve")
        _lib.check(self._lib.mm_paganin_retrieve(self._plan, x.data_ptr(), out.data_ptr(), b, _lib.stream_handle()),
                   "Paganin.retrieve")
        return res
