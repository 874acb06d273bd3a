"""PropagationLarge: the host-array interface of the Fresnel propagator.

Mirrors ``PropagationLarge`` of processing/propagation_large.py:35-348:
constructor ``PropagationLarge(n, nz, wavelength, voxelsize, distance)``,
``D(psi, j, chunks)`` with psi a host (ntheta, nz, n) or (nz, n) complex64
array or a ``BandedPinned`` and ``chunks = [CHUNK_NZ, CHUNK_2N]``,
``psi_buffer(ntheta)`` and ``free()``.  Same validation (AssertionError with
the reference's messages, processing/propagation_large.py:216-223) and
ownership: ``psi_buffer`` and ``D`` hand out cached ``BandedPinned`` buffers
banded along nz (at least 8 bands, processing/propagation_large.py:165-200),
the output overwritten by the next call.  ``DT`` (new) applies the exact
adjoint with the same interface.

The reference streams strips through a 40 GB card; here the frames move to
HBM in batches bounded by ``gpu_budget_bytes`` (several batches go through a
``StreamPipe``, copies overlapping the propagation) and the separable
two-pass kernels of libmosaic_hip.so run on them.
"""
from __future__ import annotations

import numpy as np
import torch

from model_mosaic_amd._carriers import require_gpu
from model_mosaic_amd.processing.pipeline import BandedPinned, StreamPipe, pick_n_bands
from model_mosaic_amd.processing.propagation import Propagation

DEFAULT_GPU_BUDGET = 96 << 30
OUT_N_BANDS = 8  # minimum bands of the psi and output buffers (processing/propagation_large.py:32)


def check_prop_chunks(nz: int, n: int, psi_shape, chunks) -> None:
    """Shape and [CHUNK_NZ, CHUNK_2N] validation of processing/propagation_large.py:215-223."""
    _, nz_in, n_in = psi_shape
    assert nz_in == nz and n_in == n, f"psi shape (…, {nz_in}, {n_in}) mismatches propagator ({nz}, {n})"
    chunk_nz, chunk_2n = (int(c) for c in chunks)
    assert chunk_nz >= 1 and nz % chunk_nz == 0, f"CHUNK_NZ={chunk_nz} must divide nz={nz}"
    assert chunk_2n >= 1 and (2 * n) % chunk_2n == 0, f"CHUNK_2N={chunk_2n} must divide 2n={2 * n}"


class PropagationLarge:
    """Fresnel forward propagator for host-resident frame stacks."""

    def __init__(self, n, nz, wavelength, voxelsize, distance, gpu_budget_bytes: int = DEFAULT_GPU_BUDGET,
                 pad: str = "symmetric"):
        require_gpu()
        self.n = int(n)
        self.nz = int(nz)
        self.wavelength = float(wavelength)
        self.voxelsize = float(voxelsize)
        self.distance = np.atleast_1d(np.asarray(distance, dtype=np.float64))
        self.gpu_budget_bytes = int(gpu_budget_bytes)
        self.pad = pad
        self._prop: Propagation | None = None
        self._out = None
        self._psi = None

    def _batch(self, ntheta: int) -> int:
        per = 4 * self.nz * self.n * 8  # input and output frame, two of each when streamed
        return int(max(1, min(ntheta, self.gpu_budget_bytes // per)))

    def _get_prop(self, b: int) -> Propagation:
        if self._prop is None or self._prop.ntheta < b:
            if self._prop is not None:
                self._prop.free()
            self._prop = Propagation(self.n, self.nz, b, len(self.distance), self.wavelength, self.voxelsize,
                                     self.distance, pad=self.pad)
        return self._prop

    @staticmethod
    def _banded(shape):
        return BandedPinned(shape, np.complex64, pick_n_bands(shape, np.complex64, 1, OUT_N_BANDS), band_axis=1)

    def psi_buffer(self, ntheta):
        """Cached (ntheta, nz, n) complex64 ``BandedPinned`` callers may fill before ``D``, band by band
        (processing/propagation_large.py:177-200; contents undefined on entry)."""
        shape = (int(ntheta), self.nz, self.n)
        if self._psi is None or self._psi.shape != shape:
            self._psi = self._banded(shape)
        return self._psi

    def _get_out(self, ntheta):
        shape = (ntheta, self.nz, self.n)
        if self._out is None or self._out.shape != shape:
            self._out = self._banded(shape)
        return self._out

    def free(self):
        if self._prop is not None:
            self._prop.free()
        self._prop = None
        self._out = self._psi = None

    def _run(self, psi, j, chunks, adj: bool):
        banded = isinstance(psi, BandedPinned)
        if banded:
            if psi.ndim != 3:
                raise ValueError("PropagationLarge: a BandedPinned psi must be (ntheta, nz, n)")
            shape, added_dim = psi.shape, False
        else:
            psi_t = psi if isinstance(psi, torch.Tensor) else torch.as_tensor(np.asarray(psi))
            added_dim = psi_t.ndim == 2
            if added_dim:
                psi_t = psi_t[None]
            shape = tuple(psi_t.shape)
        ntheta = shape[0]
        check_prop_chunks(self.nz, self.n, shape, chunks)
        b = self._batch(ntheta)
        prop = self._get_prop(b)
        out = self._get_out(ntheta)
        op = prop.DT if adj else prop.D
        nb = (ntheta + b - 1) // b
        if nb == 1:
            if banded:
                x = torch.empty(shape, dtype=torch.complex64, device=prop.device)
                psi.copy_to_gpu(x)
            else:
                x = psi_t.to(prop.device, dtype=torch.complex64).contiguous()
            out.copy_from_gpu(op(x, j))
            torch.cuda.synchronize(prop.device)
        else:  # frame ranges streamed: uploads and downloads overlap the propagation of the range between
            bshape = (b, self.nz, self.n)
            pipe = StreamPipe(prop.device, bshape, torch.complex64, bshape, torch.complex64)

            def load(k, pin):
                t0, t1 = k * b, min(ntheta, (k + 1) * b)
                if banded:
                    psi.copy_to(pin[: t1 - t0].numpy(), (slice(t0, t1),))
                else:
                    pin[: t1 - t0].copy_(psi_t[t0:t1])
                return pin[: t1 - t0]

            def compute(k, x, res):
                return op(x, j, out=res[: x.shape[0]])

            def store(k, r):
                out.copy_from(r.numpy(), (slice(k * b, k * b + r.shape[0]),))

            pipe.run(nb, load, compute, store)
        return out[0] if added_dim else out

    def D(self, psi, j, chunks):
        """Forward Fresnel propagation with the j-th distance (propagation_large.py:202)."""
        return self._run(psi, j, chunks, False)

    def DT(self, psi, j, chunks):
        """Adjoint of ``D`` (new; the reference provides only D)."""
        return self._run(psi, j, chunks, True)
