"""Differentiable two-physics forward model with a hand-written backward.

    L(delta, beta) = sum( |D_j( exp(i k R delta - k R beta) )|^2 - data )^2

R is the USFFT Radon transform of ``TomoReal`` (processing/tomo.py:121) and
D_j the Fresnel propagator of ``Propagation`` (processing/propagation.py:47);
the psi construction follows step3_propagation.py:94-103 with separate phase
(delta) and attenuation (beta) volumes.  The reference has no adjoint of D and
no autograd; here ``misfit`` is a ``torch.autograd.Function`` whose backward
runs DT and RT in libmosaic_hip.so (mm_model_bwd) and recomputes psi instead
of storing it.  The loss is accumulated in FP64.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, require_gpu
from model_mosaic_amd.processing.propagation import Propagation
from model_mosaic_amd.processing.tomo import TomoReal


def physical_k(wavelength: float, voxelsize: float, n: int, ntheta: int) -> float:
    """k such that k * R(delta) is the physical phase (2 pi / lambda) * voxelsize * sum(delta).

    R(x) equals the voxel line sum divided by sqrt(n * ntheta) (the phi scale of
    processing/tomo.py:59), hence k = 2 pi voxelsize sqrt(n ntheta) / lambda."""
    return 2 * math.pi / wavelength * voxelsize * math.sqrt(n * ntheta)


def reference_k(phase_scale: float, n: int, ntheta: int) -> float:
    """k of the reference pipeline: phase = p * PHASE_SCALE / sqrt(N / NTHETA) (step3_propagation.py:82,182)."""
    return float(np.float32(phase_scale / float(np.float32(np.sqrt(n / ntheta)))))


class ForwardModel:
    """Owns a Radon plan, a Fresnel plan (one distance) and the model workspace."""

    def __init__(self, n, nz, theta, wavelength, voxelsize, distance, k: float | None = None,
                 pad: str = "symmetric"):
        require_gpu()
        self._lib = _lib.load()
        self.tomo = TomoReal(n, nz, theta)
        self.ntheta = self.tomo.ntheta
        self.n, self.nz = int(n), int(nz)
        dist = np.atleast_1d(np.asarray(distance, dtype=np.float64))
        self.prop = Propagation(n, nz, self.ntheta, len(dist), wavelength, voxelsize, dist, pad=pad)
        self.prop_pad = pad
        self.k = float(k) if k is not None else physical_k(wavelength, voxelsize, n, self.ntheta)
        self.device = self.tomo.device
        h = C.c_void_p()
        _lib.check(self._lib.mm_model_create(C.byref(h), self.tomo._plan, self.prop._plan, self.nz,
                                             C.c_float(self.k)), "ForwardModel")
        self._h = h

    def free(self):
        if getattr(self, "_h", None):
            torch.cuda.synchronize(self.device)
            self._lib.mm_model_destroy(self._h)
            self._h = None
        if getattr(self, "prop", None) is not None:
            self.prop.free()
        if getattr(self, "tomo", None) is not None:
            self.tomo.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _inputs(self, delta, beta, data):
        d = as_device_tensor(delta, torch.float32, self.device)
        b = as_device_tensor(beta, torch.float32, self.device)
        y = as_device_tensor(data, torch.float32, self.device)
        if tuple(d.shape) != (self.nz, self.n, self.n) or tuple(b.shape) != (self.nz, self.n, self.n):
            raise ValueError(f"ForwardModel: delta/beta must be ({self.nz}, {self.n}, {self.n})")
        if tuple(y.shape) != (self.ntheta, self.nz, self.n):
            raise ValueError(f"ForwardModel: data must be ({self.ntheta}, {self.nz}, {self.n})")
        return d, b, y

    def loss(self, delta, beta, data, j: int = 0) -> torch.Tensor:
        """FP64 scalar loss on the device."""
        return self._forward(delta, beta, data, j)[0]

    def _forward(self, delta, beta, data, j: int = 0):
        """Loss and the token of the p, b, u it leaves in the workspace (see grad)."""
        d, b, y = self._inputs(delta, beta, data)
        out = torch.empty((), dtype=torch.float64, device=self.device)
        token = C.c_ulonglong(0)
        _lib.check(self._lib.mm_model_fwd(self._h, d.data_ptr(), b.data_ptr(), y.data_ptr(), int(j), out.data_ptr(),
                                          C.byref(token), _lib.stream_handle()), "ForwardModel.loss")
        return out, int(token.value)

    def grad(self, delta, beta, data, j: int = 0, gscale: float = 1.0, token: int = 0):
        """(dL/ddelta, dL/dbeta) times gscale, float32 (nz, n, n).  ``token`` from the forward of the same
        inputs lets the backward start from that forward's sinograms and wavefield (0: run the forward)."""
        d, b, y = self._inputs(delta, beta, data)
        gd = torch.empty_like(d)
        gb = torch.empty_like(b)
        _lib.check(self._lib.mm_model_bwd(self._h, d.data_ptr(), b.data_ptr(), y.data_ptr(), int(j),
                                          C.c_float(gscale), C.c_ulonglong(int(token)), gd.data_ptr(), gb.data_ptr(),
                                          _lib.stream_handle()), "ForwardModel.grad")
        return gd, gb


class _Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, delta, beta, model: ForwardModel, data, j: int):
        ctx.model = model
        ctx.j = j
        loss, ctx.token = model._forward(delta, beta, data, j)
        # saved tensors carry their version counters: autograd refuses a backward after an in-place change
        ctx.save_for_backward(delta, beta, data)
        return loss

    @staticmethod
    def backward(ctx, g):
        delta, beta, data = ctx.saved_tensors
        gd, gb = ctx.model.grad(delta, beta, data, ctx.j, gscale=float(g), token=ctx.token)
        return gd, gb, None, None, None


def misfit(delta: torch.Tensor, beta: torch.Tensor, model: ForwardModel, data: torch.Tensor, j: int = 0):
    """Differentiable L(delta, beta); ``.backward()`` runs the hand-written adjoint chain."""
    return _Misfit.apply(delta, beta, model, data, j)
