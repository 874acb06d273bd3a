"""PlainTomo: the reference's launch structure on one MI355X, one slice per call (benchmark comparator).

R (processing/tomo.py:121-151): zero pad, phi multiply, rfft2 (rocFFT through
torch.fft), the per-sample 81-tap gather in global memory (mm_plain_gather,
processing/kernels.py:274-349), c1d, IFFT along r, c1d, real / 4.
RT (processing/tomo.py:153-188): c1d, FFT along r, c1d, the per-sample atomic
scatter into a full centred 2n x 2n complex grid (mm_plain_scatter,
processing/kernels.py:16-93), c2d, IFFT2, c2d, crop, phi, real.
Each stage is its own pass over HBM, as in the reference.  The operators of
this package (TomoReal, TomoLargeReal) never use this module; bench.py times
it beside the batched path.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import require_gpu


class PlainTomo:
    def __init__(self, n: int, theta):
        require_gpu()
        self._lib = _lib.load()
        self.n = int(n)
        self.theta_np = np.asarray(theta, dtype=np.float32)
        self.ntheta = len(self.theta_np)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.theta = torch.from_numpy(self.theta_np).to(dev)
        n, nt = self.n, self.ntheta
        mu = -math.log(1e-3) / (2 * n * n)
        t = torch.linspace(-0.5, 0.5, n + 1, device=dev, dtype=torch.float32)[:n]
        dy, dx = torch.meshgrid(t, t, indexing="ij")
        self.phi = (torch.exp(mu * n * n * (dx * dx + dy * dy)) * (1 - n % 4) / (n * math.sqrt(n * nt)))
        self.c1d = (1 - 2 * (torch.arange(1, n + 1, device=dev) % 2)).to(torch.float32)
        c = (1 - 2 * (torch.arange(1, 2 * n + 1, device=dev) % 2)).to(torch.float32)
        self.c2d = torch.outer(c, c)

    def R(self, obj: torch.Tensor) -> torch.Tensor:
        """(n, n) float32 slice -> (ntheta, n) float32."""
        n = self.n
        pad = torch.zeros((2 * n, 2 * n), dtype=torch.float32, device=self.device)
        pad[n // 2:3 * n // 2, n // 2:3 * n // 2] = obj * self.phi
        F = torch.fft.rfft2(pad).contiguous()  # torch on ROCm returns this transform column-major
        g = torch.empty((self.ntheta, n), dtype=torch.complex64, device=self.device)
        _lib.check(self._lib.mm_plain_gather(F.data_ptr(), g.data_ptr(), self.theta.data_ptr(), n, self.ntheta,
                                             _lib.stream_handle()), "PlainTomo.R")
        g = torch.fft.ifft(g * self.c1d, dim=-1) * self.c1d
        return g.real / 4

    def RT(self, sino: torch.Tensor) -> torch.Tensor:
        """(ntheta, n) float32 slice -> (n, n) float32."""
        n = self.n
        s = torch.fft.fft(sino.to(torch.complex64) * self.c1d, dim=-1) * self.c1d
        s = s.contiguous()
        G = torch.zeros((2 * n, 2 * n), dtype=torch.complex64, device=self.device)
        _lib.check(self._lib.mm_plain_scatter(s.data_ptr(), G.data_ptr(), self.theta.data_ptr(), n, self.ntheta,
                                              _lib.stream_handle()), "PlainTomo.RT")
        G = torch.fft.ifft2(G * self.c2d) * self.c2d
        return (G[n // 2:3 * n // 2, n // 2:3 * n // 2] * self.phi).real
