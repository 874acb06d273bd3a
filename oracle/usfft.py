"""NumPy restatement of the reference USFFT Radon transform (TEST INFRASTRUCTURE ONLY).

Restates, operation by operation:

* ``TomoReal.__init__`` (processing/tomo.py:30-88): epsilon, mu, m, phi, c1dfftshift.
* ``TomoReal.R`` (processing/tomo.py:121-151) with its gather kernel
  ``gather_rfft_full`` (processing/kernels.py:274-349).
* ``TomoReal.RT`` (processing/tomo.py:153-188) with its adjoint scatter
  ``gather`` dir=1 (processing/kernels.py:16-93) and ``_lazy_init_rt``
  (processing/tomo.py:90-107).
* The arithmetic of ``TomoLargeReal.R`` / ``.RT`` (processing/tomo_large.py:72-147,
  344-838): separable phi (:81-83, 408-411), ``gather_compact`` coordinates and
  clamping (processing/kernels.py:451-526), pass-4 scale 1/(4 n sqrt(n ntheta))
  (:543) and RT1 scale 1/(n sqrt(n ntheta)) (:646).

Every function takes ``dtype``: ``np.float32`` reproduces the reference's FP32
arithmetic order (the "reference restatement"; not bit-exact, because cuFFT and
the CUDA fast intrinsics round differently from NumPy), ``np.float64`` is the
same algorithm in exact arithmetic (the FP64 oracle that the product's
accuracy bars are measured against).

Parity pins (tests/test_oracle_usfft.py):
  * gather + rfft2 against the exact polar NUDFT computed by brute force in
    ``oracle.independent`` (error O(eps) = O(1e-3) by design of the m=4 window);
  * R at theta = 0 against column sums and at theta = pi/2 against reversed row
    sums (closed forms of the Fourier-slice theorem, fixes scale and axes);
  * R of smooth Gaussian blobs against their analytic line integrals at
    arbitrary angles (fixes the sign of y0 and the angle convention);
  * RT against R through the adjoint identity <R x, y> = <x, RT y> in FP64
    (to 1e-12) for real and complex sinograms;
  * FP32 restatement against the FP64 one (rel-L2 < 1e-5);
  * TomoLargeReal restatement against TomoReal restatement (identical operator).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.fft as sfft

# threads of the 2-D transforms in tomo_R / tomo_RT (results do not depend on it: each 1-D transform is
# computed the same way on whichever thread runs it)
WORKERS = 1


def usfft_params(n):
    """epsilon, mu, m of processing/tomo.py:39-42 (and tomo_large.py:73-76)."""
    eps = 1e-3
    mu = -math.log(eps) / (2 * n * n)
    m = math.ceil(2 * n / math.pi * math.sqrt(-mu * math.log(eps) + (mu * n) ** 2 / 4))
    return eps, mu, m


def _cdtype(dtype):
    return np.complex64 if dtype == np.float32 else np.complex128


def phi_2d(n, ntheta, dtype=np.float32):
    """phi of processing/tomo.py:46-49,59 (incl. the 1/(n sqrt(n ntheta)) factor)."""
    _, mu, _ = usfft_params(n)
    t = np.linspace(-0.5, 0.5, n, endpoint=False).astype(np.float32).astype(dtype)
    dx, dy = np.meshgrid(t, t)
    arg = (mu * (n * n) * (dx * dx + dy * dy)).astype(dtype)
    phi = np.exp(arg) * dtype(1 - n % 4)
    phi = phi * dtype(1.0 / (n * np.sqrt(n * ntheta)))
    return phi.astype(dtype)


def phi_1d(n, dtype=np.float32):
    """Separable phi factor and sign of processing/tomo_large.py:81-83."""
    _, mu, _ = usfft_params(n)
    t = np.linspace(-0.5, 0.5, n, endpoint=False).astype(np.float32).astype(dtype)
    phi1d = np.exp(mu * (n * n) * (t * t)).astype(dtype)
    return phi1d, dtype(1 - n % 4)


def c1dfftshift(n):
    """processing/tomo.py:51: (-1)^(r+1) as int8."""
    return (1 - 2 * (np.arange(1, n + 1) % 2)).astype(np.int8)


def c2dfftshift(n):
    """processing/tomo.py:98-99: outer((-1)^(j+1), (-1)^(j+1)) over 2n."""
    c = (1 - 2 * (np.arange(1, 2 * n + 1) % 2)).astype(np.int8)
    return np.outer(c, c)


def _polar_coords(theta, n, dtype, variant):
    """Sample coordinates (x0, y0) in cycles/pixel, shape (ntheta, n).

    variant "real":  processing/kernels.py:293-295 (x0 = (tx - cx)/n * cos theta).
    variant "large": processing/kernels.py:473-481 (r = (j - n/2)/n, x0 = cos*r,
                     clamped to [-0.5, 0.5 - 1e-5]).
    """
    th = np.asarray(theta, dtype=np.float32).astype(dtype)
    cos_t = np.cos(th).astype(dtype)[:, None]
    sin_t = np.sin(th).astype(dtype)[:, None]
    tx = np.arange(n).astype(dtype)[None, :]
    cx = dtype(n * 0.5)
    if variant == "real":
        x0 = (tx - cx) / dtype(n) * cos_t
        y0 = -(tx - cx) / dtype(n) * sin_t
    else:
        r = (tx - cx) / dtype(n)
        x0 = cos_t * r
        y0 = -sin_t * r
        lo, hi = dtype(-0.5), dtype(0.5 - 1e-5)
        x0 = np.clip(x0, lo, hi)
        y0 = np.clip(y0, lo, hi)
    return x0.astype(dtype), y0.astype(dtype)


def _window(x0, y0, n, dtype, variant):
    """Per-sample tap rows/cols and separable Gaussian factors.

    Returns ell0 (ntheta, n, L), ell1 (ntheta, n, L), ex (…, L), ey (…, L) with
    L = 2m+1 and weight(i1, i0) = ey[i1] * ex[i0] in variant "real"
    (processing/kernels.py:306-315: ey carries coeff0) and
    coeff0*exp(coeff1*(w0^2+w1^2)) in variant "large" (kernels.py:502).
    """
    _, mu, m = usfft_params(n)
    L = 2 * m + 1
    mu_c = dtype(np.float32(mu)) if dtype == np.float32 else dtype(mu)
    pi_c = dtype(np.pi)
    coeff0 = pi_c / mu_c
    coeff1 = -(pi_c * pi_c) / mu_c
    twon = 2 * n
    ftwon = dtype(twon)
    base_x = np.rint(ftwon * x0).astype(np.int64) - m
    base_y = np.rint(ftwon * y0).astype(np.int64) - m
    i = np.arange(L)
    ell0 = base_x[..., None] + i
    ell1 = base_y[..., None] + i
    if variant == "real":
        inv_twon = dtype(1.0) / ftwon
        w0 = ell0.astype(dtype) * inv_twon - x0[..., None]
        w1 = ell1.astype(dtype) * inv_twon - y0[..., None]
    else:
        w0 = ell0.astype(dtype) / ftwon - x0[..., None]
        w1 = ell1.astype(dtype) / ftwon - y0[..., None]
    return ell0, ell1, w0.astype(dtype), w1.astype(dtype), coeff0, coeff1


def gather_half(fde, theta, n, dtype=np.float32, variant="real"):
    """Polar gather from the rfft2 half spectrum.

    fde: (nz, 2n, n+1) complex, raw fftfreq order (rfft2 output).
    Returns g (ntheta, nz, n) complex = the value written by
    ``gather_rfft_full`` (processing/kernels.py:274-349) or
    ``gather_compact`` (processing/kernels.py:451-526) for variant "large".
    """
    cdt = _cdtype(dtype)
    fde = np.asarray(fde, dtype=cdt)
    nz = fde.shape[0]
    twon = 2 * n
    x0, y0 = _polar_coords(theta, n, dtype, variant)
    ell0, ell1, w0, w1, coeff0, coeff1 = _window(x0, y0, n, dtype, variant)
    L = ell0.shape[-1]
    ntheta = x0.shape[0]
    g = np.zeros((nz, ntheta, n), dtype=cdt)
    for i1 in range(L):
        l1 = ell1[..., i1]
        ky = ((l1 % twon) + twon) % twon
        if variant == "real":
            ey = coeff0 * np.exp(coeff1 * w1[..., i1] * w1[..., i1])
        for i0 in range(L):
            l0 = ell0[..., i0]
            k0 = ((l0 % twon) + twon) % twon
            if variant == "real":
                w = ey * np.exp(coeff1 * w0[..., i0] * w0[..., i0])
            else:
                w = coeff0 * np.exp(coeff1 * (w0[..., i0] * w0[..., i0] + w1[..., i1] * w1[..., i1]))
            w = w.astype(dtype)
            sw = np.where((k0 + ky) & 1, -w, w).astype(dtype)
            refl = k0 > n
            rk_x = np.where(refl, twon - k0, k0)
            rk_y = np.where(refl, (twon - ky) % twon, ky)
            v = fde[:, rk_y, rk_x]
            v = np.where(refl[None], np.conj(v), v)
            g += sw[None] * v
    g = g / dtype(n)
    return np.ascontiguousarray(np.transpose(g, (1, 0, 2))).astype(cdt)


def scatter_full(g, theta, n, dtype=np.float32, angle_block=128):
    """Adjoint scatter of ``gather`` dir=1 (processing/kernels.py:16-93).

    g: (ntheta, nz, n) complex samples.  Returns the centred full grid
    (nz, 2n, 2n) complex with index (n + ell + 2n) % 2n (kernels.py:63,70).
    Accumulation is exact (float64 bincount) before rounding to ``dtype``; the
    reference's float atomics have no defined order.  Angles are processed in
    blocks of ``angle_block`` (the sum is the same; only host memory changes).
    """
    cdt = _cdtype(dtype)
    g = np.asarray(g, dtype=cdt)
    ntheta, nz, _ = g.shape
    twon = 2 * n
    th_all = np.asarray(theta)
    out = np.zeros((nz, twon * twon), dtype=np.complex128)
    for a0 in range(0, ntheta, angle_block):
        a1 = min(ntheta, a0 + angle_block)
        x0, y0 = _polar_coords(th_all[a0:a1], n, dtype, "real")
        ell0, ell1, w0, w1, coeff0, coeff1 = _window(x0, y0, n, dtype, "real")
        ex = np.exp(coeff1 * w0 * w0).astype(dtype)            # kernels.py:52-56
        ey = (coeff0 * np.exp(coeff1 * w1 * w1)).astype(dtype)  # kernels.py:62
        w = (ex[..., None, :] * ey[..., :, None]).astype(dtype)  # [..., i1, i0]
        fy = (n + ell1 + twon) % twon
        fx = (n + ell0 + twon) % twon
        fidx = (fy[..., :, None] * twon + fx[..., None, :]).reshape(-1)
        wf = w.reshape(a1 - a0, n, -1)
        for z in range(nz):
            c = (wf * g[a0:a1, z, :, None]).astype(cdt).reshape(-1)
            out[z] += np.bincount(fidx, weights=c.real.astype(np.float64), minlength=twon * twon)
            out[z] += 1j * np.bincount(fidx, weights=c.imag.astype(np.float64), minlength=twon * twon)
    return out.reshape(nz, twon, twon).astype(cdt)


def tomo_R(obj, theta, dtype=np.float32):
    """``TomoReal.R`` (processing/tomo.py:121-151).  obj (nz, n, n) -> (ntheta, nz, n)."""
    obj = np.asarray(obj)
    nz, n, _ = obj.shape
    ntheta = len(theta)
    cdt = _cdtype(dtype)
    phi = phi_2d(n, ntheta, dtype)
    padded = np.zeros((nz, 2 * n, 2 * n), dtype=dtype)
    padded[:, n // 2:3 * n // 2, n // 2:3 * n // 2] = obj.astype(dtype) * phi
    fde = sfft.rfft2(padded, axes=(-2, -1), workers=WORKERS).astype(cdt)
    sino = gather_half(fde, theta, n, dtype, "real")
    c1d = c1dfftshift(n)
    sino = (sino * c1d).astype(cdt)
    sino = sfft.ifft(sino, axis=-1).astype(cdt)
    sino = (sino * c1d).astype(cdt)
    return (sino.real / dtype(4)).astype(dtype)


def tomo_RT(data, theta, n=None, dtype=np.float32):
    """``TomoReal.RT`` (processing/tomo.py:153-188).

    data (ntheta, nz, n) real or complex -> (nz, n, n), real for real input.
    """
    data = np.asarray(data)
    ntheta, nz, n = data.shape
    cdt = _cdtype(dtype)
    is_real = not np.iscomplexobj(data)
    c1d = c1dfftshift(n)
    s = (data.astype(cdt if not is_real else dtype) * c1d).astype(cdt)
    s = sfft.fft(s, axis=-1).astype(cdt)
    s = (s * c1d).astype(cdt)
    fde = scatter_full(s, theta, n, dtype)
    c2d = c2dfftshift(n)
    fde = (fde * c2d).astype(cdt)
    fde = sfft.ifft2(fde, axes=(-2, -1), workers=WORKERS).astype(cdt)
    fde = (fde * c2d).astype(cdt)
    phi = phi_2d(n, ntheta, dtype)
    res = (fde[:, n // 2:3 * n // 2, n // 2:3 * n // 2] * phi).astype(cdt)
    return res.real.astype(dtype) if is_real else res


def tomo_large_R(obj, theta, dtype=np.float32):
    """Arithmetic of ``TomoLargeReal.R`` (processing/tomo_large.py:344-580) in one piece.

    Separable phi without the 1/(n sqrt(n ntheta)) factor (pass1 :408-411),
    rfft along x then full FFT along y (pass1/pass2), compact gather with
    clamped coordinates (pass3, kernels.py:451-526), pass4 scale
    1/(4 n sqrt(n ntheta)) (:543) and ``.real``.  Chunking does not change the
    arithmetic, so no chunk arguments are needed here.
    """
    obj = np.asarray(obj)
    nz, n, _ = obj.shape
    ntheta = len(theta)
    cdt = _cdtype(dtype)
    phi1d, phi_scale = phi_1d(n, dtype)
    padded = np.zeros((nz, n, 2 * n), dtype=dtype)
    padded[:, :, n // 2:n // 2 + n] = phi_scale * phi1d[None, :, None] * phi1d[None, None, :] * obj.astype(dtype)
    rows = sfft.rfft(padded, axis=-1).astype(cdt)
    fde = np.zeros((nz, 2 * n, n + 1), dtype=cdt)
    fde[:, n // 2:n // 2 + n, :] = rows
    fde = sfft.fft(fde, axis=1).astype(cdt)
    sino = gather_half(fde, theta, n, dtype, "large")
    c1d = c1dfftshift(n)
    scale = dtype(1.0 / (4 * n * np.sqrt(n * ntheta)))
    sino = (sino * c1d).astype(cdt)
    sino = sfft.ifft(sino, axis=-1).astype(cdt)
    sino = (sino * c1d).astype(cdt)
    sino = (sino * scale).astype(cdt)
    return sino.real.astype(dtype)


def tomo_large_RT(sino, theta, dtype=np.float32):
    """Arithmetic of ``TomoLargeReal.RT`` (processing/tomo_large.py:582-838) in one piece.

    RT1 scale 1/(n sqrt(n ntheta)) (:646), full-grid scatter (the union of the
    ky bands of ``scatter_compact``), c2d-masked IFFT along y then x, crop,
    separable phi (:817-828); real output for real input.
    """
    sino = np.asarray(sino)
    ntheta, nz, n = sino.shape
    cdt = _cdtype(dtype)
    is_real = not np.iscomplexobj(sino)
    c1d = c1dfftshift(n)
    scale = dtype(1.0 / (n * np.sqrt(n * ntheta)))
    s = (sino.astype(cdt) * c1d).astype(cdt)
    s = sfft.fft(s, axis=-1).astype(cdt)
    s = (s * c1d).astype(cdt)
    s = (s * scale).astype(cdt)
    fde = scatter_full(s, theta, n, dtype)
    c2d = c2dfftshift(n)
    fde = (fde * c2d).astype(cdt)
    fde = sfft.ifft(fde, axis=1).astype(cdt)
    fde = sfft.ifft(fde, axis=2).astype(cdt)
    fde = (fde * c2d).astype(cdt)
    phi1d, phi_scale = phi_1d(n, dtype)
    cropped = fde[:, n // 2:n // 2 + n, n // 2:n // 2 + n]
    res = (phi_scale * phi1d[None, :, None] * phi1d[None, None, :] * cropped).astype(cdt)
    return res.real.astype(dtype) if is_real else res
