"""NumPy restatement of the reference Paganin phase retrieval (TEST INFRASTRUCTURE ONLY).

Restates processing/paganin.py:
  * ``Paganin.__init__`` filter (:47-56): fy = fftfreq(nz, d), fx =
    rfftfreq(n, d) (float32 in the reference), w2 = fy^2 + fx^2,
    filt = alpha / (wavelength * distance * w2 / (4 pi) + alpha);
  * ``Paganin.retrieve`` (:70-92): rfft2 over (nz, n), * filt, irfft2 with
    s = (nz, n), -log(clip(., 1e-6)).
``PaganinLarge.retrieve`` (processing/paganin_large.py:115-284) computes the
same map in x-strips and z-strips; its chunking does not change the result.

Parity pins (tests/test_oracle_paganin.py): H(0, 0) = 1, so a constant frame
maps to -log(constant); distance 0 makes H = 1 (the map is -log(clip(I)));
a single Fourier mode I = c + e cos(2 pi (p y / nz + q x / n)) maps to
-log(c + e H(p, q) cos(...)) with H from the closed form at that mode (a
wrong sign, factor or frequency scale in the filter fails it); the clip at
1e-6 for non-positive inputs.
"""
from __future__ import annotations

import numpy as np


def filt(n, nz, wavelength, voxelsize, distance, alpha, dtype=np.float64):
    """(nz, n // 2 + 1) filter on the rfft2 grid (processing/paganin.py:47-56)."""
    fy = np.fft.fftfreq(nz, d=voxelsize).astype(dtype)
    fx = np.fft.rfftfreq(n, d=voxelsize).astype(dtype)
    w2 = fy[:, None] ** 2 + fx[None, :] ** 2
    dt = np.dtype(dtype).type
    return (dt(alpha) / (dt(wavelength * distance) * w2 / dt(4.0 * np.pi) + dt(alpha))).astype(dtype)


def retrieve(intensity, wavelength, voxelsize, distance, alpha, dtype=np.float64):
    """processing/paganin.py:70-92 on (b, nz, n) or (nz, n) intensities."""
    x = np.asarray(intensity, dtype=dtype)
    nz, n = x.shape[-2:]
    h = filt(n, nz, wavelength, voxelsize, distance, alpha, dtype)
    y = np.fft.irfft2(np.fft.rfft2(x, axes=(-2, -1)) * h, s=(nz, n), axes=(-2, -1))
    return -np.log(np.clip(y, 1e-6, None)).astype(dtype)


def mode_gain(p, q, n, nz, wavelength, voxelsize, distance, alpha):
    """Closed form of H at the Fourier mode (ky, kx) = (p / (nz d), q / (n d))."""
    k2 = (p / (nz * voxelsize)) ** 2 + (q / (n * voxelsize)) ** 2
    return alpha / (wavelength * distance * k2 / (4 * np.pi) + alpha)
