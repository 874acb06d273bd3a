"""TomoLargeReal: the host-array interface of the USFFT Radon transform.

Mirrors ``TomoLargeReal`` of processing/tomo_large.py:46-838: constructor
``TomoLargeReal(n, theta, chunk_xy)``, ``R(obj, chunks)`` / ``RT(sino, chunks)``
on host (numpy or pinned) arrays with ``chunks = [CHUNK_N, CHUNK_THETA,
CHUNK_XY]``, ``obj_buffer(nz)`` and ``free()``, and the same ownership rules:
``R`` returns a view into an internal pinned buffer that stays valid until the
next ``R`` call (processing/tomo_large.py:545-553); ``RT`` returns the internal
pinned object buffer.

On the reference the chunks bound GPU memory of a 40 GB card.  With 288 GB of
HBM the MI355X path is device resident: the whole z range is moved once per
call and computed by the same kernels as ``TomoReal``; when the volume
exceeds ``gpu_budget_bytes`` it is streamed in z slabs through a
``StreamPipe`` (upload of slab k+1 and download of slab k-1 overlap the
transform of slab k).  The chunk arguments are still
validated as in the reference (AssertionError for a CHUNK_N that does not
divide n, processing/tomo_large.py:430), and a CHUNK_THETA that does not divide
ntheta is rejected too (the reference would silently skip the tail angles of
its pass 4, processing/tomo_large.py:577).

Arithmetic note: the reference's compact gather clamps the polar coordinates to
[-0.5, 0.5 - 1e-5] (processing/kernels.py:478-481), which moves the r = 0
sample at theta ~ pi/2 by 1e-5 cycles/pixel.  This class uses the unclamped
``TomoReal`` geometry; the two differ by ~3e-8 relative L2, far inside the
reference's own 1e-4 Real-vs-Large tolerance (tests/test_large_variants.py:73,77).
"""
from __future__ import annotations

import numpy as np
import torch

from model_mosaic_amd._carriers import pinned_empty, require_gpu
from model_mosaic_amd.processing.pipeline import StreamPipe
from model_mosaic_amd.processing.tomo import TomoReal

DEFAULT_GPU_BUDGET = 96 << 30  # bytes of HBM one call may hold


def check_tomo_chunks(n: int, ntheta: int, chunks) -> None:
    """[CHUNK_N, CHUNK_THETA, CHUNK_XY] validation (processing/tomo_large.py:430 and its pass loops)."""
    if len(chunks) != 3:
        raise AssertionError(f"chunks must be [CHUNK_N, CHUNK_THETA, CHUNK_XY], got {list(chunks)}")
    chunk_n, chunk_theta, chunk_xy = (int(c) for c in chunks)
    assert chunk_n >= 1 and n % chunk_n == 0, f"CHUNK_N={chunk_n} must divide n={n} (rfft x-axis strips)"
    assert chunk_theta >= 1 and ntheta % chunk_theta == 0, f"CHUNK_THETA={chunk_theta} must divide ntheta={ntheta}"
    assert chunk_xy >= 1, f"CHUNK_XY={chunk_xy} must be >= 1"


class TomoLargeReal:
    """Radon transform via USFFT for host-resident volumes."""

    def __init__(self, n, theta, chunk_xy, gpu_budget_bytes: int = DEFAULT_GPU_BUDGET, fp64_tiles: bool = False):
        require_gpu()
        self.n = int(n)
        self.theta = np.asarray(theta).astype(np.float32)
        self.ntheta = len(self.theta)
        if int(chunk_xy) < 1:
            raise AssertionError(f"CHUNK_XY={chunk_xy} must be >= 1")
        self.chunk_xy = int(chunk_xy)
        self.gpu_budget_bytes = int(gpu_budget_bytes)
        self.fp64_tiles = fp64_tiles
        self._tomo: TomoReal | None = None
        self._sino = None    # pinned (ntheta, nz, n) float32 (torch)
        self._obj = None     # pinned (nz, n, n) float32 for obj_buffer
        self._obj_rt = None  # pinned (nz, n, n) float32 / complex64

    # ---------- sizing ------------------------------------------------------
    def _slice_bytes(self, cplx: bool) -> int:
        n, nt = self.n, self.ntheta
        rows = n * (n + 4) * 8 if n > 16384 else 0           # widths above 16384: the buffer of object rows
        per = 2 * n * (n + 2) * 8 + rows + nt * (n // 2 + 1) * 8   # plan workspace per slice
        per += 2 * (4 * n * n + 4 * nt * n)                  # obj + sino slabs, two of each when streamed
        if cplx:
            per += 2 * n * (n + 2) * 8 + rows + nt * (n // 2 + 1) * 8 + 2 * (4 * n * n + 4 * nt * n)
        return per

    def _zbatch(self, nz: int, cplx: bool = False) -> int:
        return int(max(1, min(nz, self.gpu_budget_bytes // self._slice_bytes(cplx))))

    def _get_tomo(self, zb: int) -> TomoReal:
        if self._tomo is None or self._tomo._nz < zb:
            if self._tomo is not None:
                self._tomo.free()
            self._tomo = TomoReal(self.n, zb, self.theta, fp64_tiles=self.fp64_tiles)
        return self._tomo

    def _check_chunks(self, chunks):
        check_tomo_chunks(self.n, self.ntheta, chunks)

    # ---------- buffers -----------------------------------------------------
    def obj_buffer(self, nz):
        """Pinned (nz, n, n) float32 buffer a caller may fill before ``R`` (tomo_large.py:327)."""
        shape = (int(nz), self.n, self.n)
        if self._obj is None or tuple(self._obj.shape) != shape:
            self._obj = pinned_empty(shape, torch.float32)
        return self._obj.numpy()

    def _get_sino(self, nz):
        shape = (self.ntheta, nz, self.n)
        if self._sino is None or tuple(self._sino.shape) != shape:
            self._sino = pinned_empty(shape, torch.float32)
        return self._sino

    def _get_obj_rt(self, nz, dtype):
        shape = (nz, self.n, self.n)
        if self._obj_rt is None or tuple(self._obj_rt.shape) != shape or self._obj_rt.dtype != dtype:
            self._obj_rt = pinned_empty(shape, dtype)
        return self._obj_rt

    def free(self):
        """Releases the device plan and the pinned buffers (tomo_large.py:260)."""
        if self._tomo is not None:
            self._tomo.free()
        self._tomo = None
        self._sino = self._obj = self._obj_rt = None

    # ---------- operators ---------------------------------------------------
    def R(self, obj, chunks):
        """(nz, n, n) float32 host obj -> (ntheta, nz, n) float32 host sino (view, valid until next R)."""
        self._check_chunks(chunks)
        obj_t = torch.as_tensor(np.asarray(obj) if not isinstance(obj, torch.Tensor) else obj)
        if obj_t.ndim != 3 or obj_t.shape[1] != self.n or obj_t.shape[2] != self.n:
            raise ValueError(f"TomoLargeReal.R: obj shape {tuple(obj_t.shape)} != (nz, {self.n}, {self.n})")
        nz = obj_t.shape[0]
        zb = self._zbatch(nz)
        tomo = self._get_tomo(zb)
        sino = self._get_sino(nz)
        nb = (nz + zb - 1) // zb
        if nb == 1:
            sino.copy_(tomo.R(obj_t.to(tomo.device, dtype=torch.float32)))
            torch.cuda.synchronize(tomo.device)
            return sino.numpy()
        # z slabs streamed: upload of slab k+1 and download of slab k-1 overlap the transform of slab k
        pipe = StreamPipe(tomo.device, (zb, self.n, self.n), torch.float32, (self.ntheta, zb, self.n), torch.float32)

        def load(k, pin):
            z0, z1 = k * zb, min(nz, (k + 1) * zb)
            v = pin[: z1 - z0]
            v.copy_(obj_t[z0:z1])
            return v

        def compute(k, x, out):
            return tomo.R(x, out=out.view(-1)[: self.ntheta * x.shape[0] * self.n].view(self.ntheta, x.shape[0], self.n))

        def store(k, res):
            sino[:, k * zb: k * zb + res.shape[1]].copy_(res)

        pipe.run(nb, load, compute, store)
        return sino.numpy()

    def RT(self, sino, chunks):
        """(ntheta, nz, n) host sino (float32 or complex64) -> pinned (nz, n, n) of the same kind."""
        self._check_chunks(chunks)
        s_np = sino if isinstance(sino, torch.Tensor) else torch.as_tensor(np.asarray(sino))
        if s_np.ndim != 3 or s_np.shape[0] != self.ntheta or s_np.shape[2] != self.n:
            raise ValueError(f"TomoLargeReal.RT: sino shape {tuple(s_np.shape)} != ({self.ntheta}, nz, {self.n})")
        is_real = s_np.dtype == torch.float32
        dt = torch.float32 if is_real else torch.complex64
        nz = s_np.shape[1]
        zb = self._zbatch(nz, cplx=not is_real)
        tomo = self._get_tomo(zb)
        out = self._get_obj_rt(nz, dt)
        nb = (nz + zb - 1) // zb
        if nb == 1:
            out.copy_(tomo.RT(s_np.to(tomo.device, dtype=dt).contiguous()))
            torch.cuda.synchronize(tomo.device)
            return out.numpy()
        pipe = StreamPipe(tomo.device, (self.ntheta, zb, self.n), dt, (zb, self.n, self.n), dt)

        def load(k, pin):
            z0, z1 = k * zb, min(nz, (k + 1) * zb)
            v = pin.view(-1)[: self.ntheta * (z1 - z0) * self.n].view(self.ntheta, z1 - z0, self.n)
            v.copy_(s_np[:, z0:z1])
            return v

        def compute(k, y, res):
            return tomo.RT(y, out=res.view(-1)[: y.shape[1] * self.n * self.n].view(y.shape[1], self.n, self.n))

        def store(k, res):
            out[k * zb: k * zb + res.shape[0]].copy_(res)

        pipe.run(nb, load, compute, store)
        return out.numpy()
