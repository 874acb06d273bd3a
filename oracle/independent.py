"""Oracles that share no code with the restatement (TEST INFRASTRUCTURE ONLY).

* ``nudft_polar``: brute-force FP64 non-uniform DFT of an image at the polar
  sample frequencies of the USFFT gather.  The USFFT with m = 4 approximates it
  to O(eps) = O(1e-3) (processing/tomo.py:39-42 designs the window for
  eps = 1e-3).  Derivation of the constant: the gather sums
  (pi/mu) exp(-pi^2 |xi - l/2n|^2 / mu) F_c(l) over the centred 2n-grid FFT of
  phi*obj; by the Gaussian Fourier pair this is 4 n^2 times the NUDFT of obj,
  divided by n in the kernel (processing/kernels.py:344) and scaled by
  (1 - n%4)/(n sqrt(n ntheta)) through phi (processing/tomo.py:49,59), so
  g = 4 (1 - n%4) NUDFT(obj)(x0, y0) / sqrt(n ntheta).
* ``radon_from_nudft``: the same c1dfftshift / IFFT / .real/4 tail as
  ``TomoReal.R`` applied to the exact NUDFT samples.
* ``gaussian_blob`` / ``gaussian_blob_projection``: a smooth blob and its
  analytic line integral p(s) = A sqrt(2 pi) sigma exp(-(s - s_c)^2 / 2 sigma^2),
  s = x cos(theta) - y sin(theta), x, y measured from pixel n/2 (the Fourier
  slice along (cos theta, -sin theta) of processing/kernels.py:294-295).
* ``cosine_grating`` / ``cosine_grating_propagated``: a field that symmetric
  padding keeps periodic on the padded grid, so Fresnel propagation multiplies
  each of its harmonics by exp(-i pi lambda z f^2) exactly.
"""
from __future__ import annotations

import numpy as np

from oracle.usfft import c1dfftshift


def nudft_polar(obj, theta):
    """Exact polar NUDFT in FP64: (nz, n, n) -> (ntheta, nz, n) complex128."""
    obj = np.asarray(obj, dtype=np.float64)
    nz, n, _ = obj.shape
    th = np.asarray(theta, dtype=np.float32).astype(np.float64)
    xs = np.arange(n) - n / 2.0
    r = (np.arange(n) - n / 2.0) / n
    out = np.empty((len(th), nz, n), dtype=np.complex128)
    for k, t in enumerate(th):
        xi = r * np.cos(t)
        eta = -r * np.sin(t)
        ex = np.exp(-2j * np.pi * xi[:, None] * xs[None, :])    # (n_r, n_x)
        ey = np.exp(-2j * np.pi * eta[:, None] * xs[None, :])   # (n_r, n_y)
        # F[z, r] = sum_{y,x} obj[z,y,x] ey[r,y] ex[r,x]: the x sum as one matrix product per slice
        for z in range(nz):
            tmp = obj[z] @ ex.T                                  # (n_y, n_r)
            out[k, z] = np.sum(tmp * ey.T, axis=0)
    return out


def radon_from_nudft(obj, theta):
    """FP64 Radon through exact NUDFT samples; the USFFT approximates this to O(eps)."""
    obj = np.asarray(obj)
    nz, n, _ = obj.shape
    ntheta = len(theta)
    g = 4.0 * (1 - n % 4) * nudft_polar(obj, theta) / np.sqrt(n * ntheta)
    c1d = c1dfftshift(n)
    s = np.fft.ifft(g * c1d, axis=-1) * c1d
    return s.real / 4.0


def gaussian_blob(n, nz, centres, sigmas, amps):
    """Sum of isotropic 2-D Gaussian blobs, same in every z slice: (nz, n, n)."""
    x = np.arange(n) - n / 2.0
    X, Y = np.meshgrid(x, x)
    img = np.zeros((n, n))
    for (xc, yc), s, a in zip(centres, sigmas, amps):
        img += a * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2 * s * s))
    return np.repeat(img[None], nz, axis=0)


def gaussian_blob_projection(n, theta, centres, sigmas, amps):
    """Analytic line integrals of ``gaussian_blob``: (ntheta, n) at s = t - n/2."""
    th = np.asarray(theta, dtype=np.float32).astype(np.float64)
    s = np.arange(n) - n / 2.0
    out = np.zeros((len(th), n))
    for (xc, yc), sg, a in zip(centres, sigmas, amps):
        sc = xc * np.cos(th) - yc * np.sin(th)
        out += a * np.sqrt(2 * np.pi) * sg * np.exp(-(s[None, :] - sc[:, None]) ** 2 / (2 * sg * sg))
    return out


def cosine_grating(nz, n, px, py, a):
    """1 + a cos(2 pi (x+1/2)/px) cos(2 pi (y+1/2)/py); px | n, py | nz keep it symmetric-pad periodic."""
    x = np.arange(n) + 0.5
    y = np.arange(nz) + 0.5
    cx = np.cos(2 * np.pi * x / px) if px else np.ones(n)
    cy = np.cos(2 * np.pi * y / py) if py else np.ones(nz)
    return 1.0 + a * cy[:, None] * cx[None, :]


def cosine_grating_propagated(nz, n, px, py, a, wavelength, voxelsize, z):
    """Closed-form Fresnel propagation of ``cosine_grating`` (complex128)."""
    f2 = 0.0
    if px:
        f2 += 1.0 / (px * voxelsize) ** 2
    if py:
        f2 += 1.0 / (py * voxelsize) ** 2
    chirp = np.exp(-1j * np.pi * wavelength * z * f2)
    base = cosine_grating(nz, n, px, py, a) - 1.0
    return 1.0 + chirp * base
