"""TomoReal: USFFT Radon transform on one MI355X, device-resident.

Mirrors ``TomoReal`` of processing/tomo.py:17-188 (same constructor
``TomoReal(n, nz, theta)``, ``R``, ``RT``, ``free``; same layouts: obj
(nz, n, n) float32, sinogram (ntheta, nz, n)).  Carriers are torch ROCm
tensors instead of cupy arrays; numpy inputs are copied to the device.
All arithmetic happens in libmosaic_hip.so (mm_radon_*).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, check_out, require_gpu


class TomoReal:
    """Radon transform R and adjoint RT via the USFFT (eps = 1e-3, m = 4, 2x oversampling).

    Parameters mirror processing/tomo.py:30: ``n`` detector width (= object
    width), ``nz`` the maximum number of slices per call, ``theta`` the
    projection angles in radians (stored as float32, like the reference).
    The adjoint scatter runs as dense 4 x 4-patch products on the matrix cores (FP32, 16 slices per
    product, so only for ``nz >= 16``; smaller plans take the per-point lists) unless one of
    ``fp64_tiles`` (tiles accumulate in FP64), ``kahan_tiles`` (FP32 with Kahan compensation) or
    ``list_tiles`` (FP32 on the vector ALUs from per-point tap lists) is set (at most one of them).
    ``gather_zp`` (1 or 2) is the number of z slices the forward gather stages
    per patch element for tiles of width <= 32 (2, the default, is the faster
    variant on gfx950; wider tiles always stage one slice); ``gather_root``
    (64 or 32) is the largest quadtree tile.  Neither changes the result.
    """

    def __init__(self, n, nz, theta, fp64_tiles: bool = False, gather_zp: int = 2, kahan_tiles: bool = False,
                 list_tiles: bool = False, gather_root: int = 64):
        require_gpu()
        self._lib = _lib.load()
        self.n = int(n)
        self._nz = int(nz)
        self.theta = np.asarray(theta).astype(np.float32)
        self.ntheta = len(self.theta)
        self.device = torch.device("cuda", torch.cuda.current_device())
        th = np.ascontiguousarray(self.theta.astype(np.float64))
        plan = C.c_void_p()
        if gather_zp not in (1, 2):
            raise ValueError(f"TomoReal: gather_zp must be 1 or 2, got {gather_zp}")
        if gather_root not in (32, 64):
            raise ValueError(f"TomoReal: gather_root must be 32 or 64, got {gather_root}")
        if int(fp64_tiles) + int(kahan_tiles) + int(list_tiles) > 1:
            raise ValueError("TomoReal: at most one of fp64_tiles, kahan_tiles, list_tiles")
        flags = ((_lib.MM_RADON_FP64_TILES if fp64_tiles else 0) | (_lib.MM_RADON_GATHER_ZP1 if gather_zp == 1 else 0)
                 | (_lib.MM_RADON_KAHAN_TILES if kahan_tiles else 0) | (_lib.MM_RADON_LIST_TILES if list_tiles else 0)
                 | (_lib.MM_RADON_GATHER_ROOT32 if gather_root == 32 else 0))
        _lib.check(self._lib.mm_radon_create(C.byref(plan), self.n, self._nz, self.ntheta,
                                             th.ctypes.data_as(C.POINTER(C.c_double)), flags),
                   "TomoReal")
        self._plan = plan

    # -- helpers -------------------------------------------------------------
    def _alive(self):
        if not self._plan:
            raise RuntimeError("TomoReal: plan was freed")

    def workspace_bytes(self) -> int:
        return int(self._lib.mm_radon_workspace_bytes(self._plan)) if self._plan else 0

    def free(self):
        """Releases the device plan (processing/tomo.py:109)."""
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            self._lib.mm_radon_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # -- operators -----------------------------------------------------------
    def R(self, obj, out: torch.Tensor | None = None) -> torch.Tensor:
        """(nz, n, n) float32 object -> (ntheta, nz, n) float32 sinogram (processing/tomo.py:121)."""
        self._alive()
        x = as_device_tensor(obj, torch.float32, self.device)
        if x.ndim != 3 or x.shape[1] != self.n or x.shape[2] != self.n:
            raise ValueError(f"TomoReal.R: obj shape {tuple(x.shape)} != (nz, {self.n}, {self.n})")
        nz = x.shape[0]
        if not 1 <= nz <= self._nz:
            raise ValueError(f"TomoReal.R: nz={nz} outside [1, {self._nz}]")
        if out is None:
            out = torch.empty((self.ntheta, nz, self.n), dtype=torch.float32, device=self.device)
        else:
            check_out(out, (self.ntheta, nz, self.n), torch.float32, self.device, "TomoReal.R")
        _lib.check(self._lib.mm_radon_fwd(self._plan, x.data_ptr(), out.data_ptr(), nz, _lib.stream_handle()),
                   "TomoReal.R")
        return out

    def RT(self, data, out: torch.Tensor | None = None) -> torch.Tensor:
        """(ntheta, nz, n) sinogram -> (nz, n, n) object (processing/tomo.py:153).

        float32 input returns the real part (float32); any other dtype is taken
        as complex64 and returns the complex result, like the reference.
        """
        self._alive()
        is_real = (data.dtype == torch.float32) if isinstance(data, torch.Tensor) else (np.asarray(data).dtype == np.float32)
        dt = torch.float32 if is_real else torch.complex64
        y = as_device_tensor(data, dt, self.device)
        if y.ndim != 3 or y.shape[0] != self.ntheta or y.shape[2] != self.n:
            raise ValueError(f"TomoReal.RT: data shape {tuple(y.shape)} != ({self.ntheta}, nz, {self.n})")
        nz = y.shape[1]
        if not 1 <= nz <= self._nz:
            raise ValueError(f"TomoReal.RT: nz={nz} outside [1, {self._nz}]")
        if out is None:
            out = torch.empty((nz, self.n, self.n), dtype=dt, device=self.device)
        else:
            check_out(out, (nz, self.n, self.n), dt, self.device, "TomoReal.RT")
        _lib.check(self._lib.mm_radon_adj(self._plan, y.data_ptr(), 0 if is_real else 1, out.data_ptr(), nz,
                                          _lib.stream_handle()), "TomoReal.RT")
        return out

    def RT_filtered(self, data, fbp, w, out: torch.Tensor | None = None) -> torch.Tensor:
        """RT of the FBP-filtered real sinogram, ``fbp.filter(data, w)`` then ``RT`` (step8_fbp.py:204-205),
        as one pass over the sinogram: the filter (length-2n form, ``FBPFilter``) runs inside the line
        transform of RT, so the filtered sinogram never goes to HBM and ``data`` is left unchanged."""
        self._alive()
        y = as_device_tensor(data, torch.float32, self.device).contiguous()
        if y.ndim != 3 or y.shape[0] != self.ntheta or y.shape[2] != self.n:
            raise ValueError(f"TomoReal.RT_filtered: data shape {tuple(y.shape)} != ({self.ntheta}, nz, {self.n})")
        if fbp.n != self.n:
            raise ValueError(f"TomoReal.RT_filtered: filter n={fbp.n} != {self.n}")
        nz = y.shape[1]
        if not 1 <= nz <= self._nz:
            raise ValueError(f"TomoReal.RT_filtered: nz={nz} outside [1, {self._nz}]")
        fbp.check_weights(w)
        G, A, B = fbp._two_n(w)
        if out is None:
            out = torch.empty((nz, self.n, self.n), dtype=torch.float32, device=self.device)
        else:
            check_out(out, (nz, self.n, self.n), torch.float32, self.device, "TomoReal.RT_filtered")
        _lib.check(self._lib.mm_radon_adj_fbp(self._plan, y.data_ptr(), G.data_ptr(), A.data_ptr(), B.data_ptr(),
                                              out.data_ptr(), nz, _lib.stream_handle()), "TomoReal.RT_filtered")
        return out
