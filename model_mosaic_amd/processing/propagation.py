"""Propagation: Fresnel near-field propagator D and its adjoint DT on one MI355X.

Mirrors ``Propagation`` of processing/propagation.py:15-64 (constructor
``Propagation(n, nz, ntheta, ndist, wavelength, voxelsize, distance)``,
``D(psi, j)`` with psi (b <= ntheta, nz, n) or (nz, n) complex64, returning a
new array).  ``DT`` is new: the exact adjoint of ``D``.  Symmetric mirror
padding to (2nz, 2n) as in processing/kernels.py:530-550; ``pad="zero"``
selects zero padding.  All arithmetic happens in libmosaic_hip.so
(mm_fresnel_*).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, check_out, require_gpu


class Propagation:
    """Fresnel forward propagator sized for (ntheta x nz x n) frames."""

    def __init__(self, n, nz, ntheta, ndist, wavelength, voxelsize, distance, pad: str = "symmetric"):
        require_gpu()
        self._lib = _lib.load()
        self.n = int(n)
        self.nz = int(nz)
        self.ntheta = int(ntheta)
        self.ndist = int(ndist)
        self.wavelength = float(wavelength)
        self.voxelsize = float(voxelsize)
        self.distance = np.atleast_1d(np.asarray(distance, dtype=np.float64))[: self.ndist].copy()
        if len(self.distance) != self.ndist:
            raise ValueError(f"Propagation: {len(self.distance)} distances for ndist={self.ndist}")
        if pad not in ("symmetric", "zero"):
            raise ValueError("Propagation: pad must be 'symmetric' or 'zero'")
        self.device = torch.device("cuda", torch.cuda.current_device())
        plan = C.c_void_p()
        d = np.ascontiguousarray(self.distance)
        _lib.check(self._lib.mm_fresnel_create(C.byref(plan), self.n, self.nz, self.ndist, self.wavelength,
                                               self.voxelsize, d.ctypes.data_as(C.POINTER(C.c_double)),
                                               0 if pad == "symmetric" else 1), "Propagation")
        self._plan = plan

    def free(self):
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            self._lib.mm_fresnel_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def _apply(self, psi, j, adj: bool, out=None):
        if not self._plan:
            raise RuntimeError("Propagation: plan was freed")
        x = as_device_tensor(psi, torch.complex64, self.device)
        added_dim = x.ndim == 2
        if added_dim:
            x = x[None]
        if x.ndim != 3 or x.shape[1] != self.nz or x.shape[2] != self.n:
            raise ValueError(f"Propagation: psi shape {tuple(x.shape)} != (b, {self.nz}, {self.n})")
        b = x.shape[0]
        if b > self.ntheta:
            raise ValueError(f"Propagation: batch {b} exceeds ntheta={self.ntheta}")
        what = "Propagation.DT" if adj else "Propagation.D"
        if out is None:
            res = out = torch.empty_like(x)
            if added_dim:
                res = out[0]
        else:
            res = check_out(out, x.shape, torch.complex64, self.device, what)
        fn = self._lib.mm_fresnel_adj if adj else self._lib.mm_fresnel_fwd
        _lib.check(fn(self._plan, x.data_ptr(), out.data_ptr(), b, int(j), _lib.stream_handle()), what)
        return res

    def D(self, psi, j, out=None) -> torch.Tensor:
        """Forward Fresnel propagation with the j-th distance (processing/propagation.py:47)."""
        return self._apply(psi, j, False, out)

    def DT(self, psi, j, out=None) -> torch.Tensor:
        """Adjoint of ``D``: <D x, y> = <x, DT y>."""
        return self._apply(psi, j, True, out)
