"""NumPy restatement of the reference Fresnel propagator (TEST INFRASTRUCTURE ONLY).

Restates:

* ``Propagation.__init__`` kernels (processing/propagation.py:22-31): FP32
  fftfreq grids, f2 = fx^2 + fy^2, fker = exp(-i pi lambda z f2) / (4 n nz).
* ``pad_fwd`` symmetric mirror padding (processing/kernels.py:530-550).
* ``Propagation.D`` (processing/propagation.py:47-64): pad, fft2, * fker,
  ifft2(norm="forward"), centre crop.
* ``PropagationLarge.D`` (processing/propagation_large.py:202-348): separable
  K_x (2n) and K_y (2nz) kernels (:69-79), x-pad/FFT/K_x, y-pad/FFT/K_y/IFFT/
  crop, x-IFFT/crop.
* ``DT``: not in the reference (propagation_large.py:15 says only D exists);
  derived here as the exact adjoint of D: zero-embed, FFT2, * conj(fker),
  IFFT2(norm="forward"), transpose of the mirror pad (fold-add).
* ``_psi_from_proj`` and ``_abs2_c64_to_f32`` (step3_propagation.py:94-110).

Parity pins (tests/test_oracle_fresnel.py): z = 0 is the identity; a constant
field is unchanged; mirror pad equals ``np.pad(mode="symmetric")``; the padded
propagator has |fker * 4 n nz| = 1; 2-D equals separable; the cosine-grating
closed form of ``oracle.independent`` (Talbot phase); DT satisfies
<D x, y> = <x, DT y> to 1e-12 in FP64; psi/abs2 against their closed forms.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft


def _cdt(dtype):
    return np.complex64 if dtype == np.float32 else np.complex128


def fresnel_kernel_2d(n, nz, wavelength, voxelsize, distance, dtype=np.float32):
    """fker of processing/propagation.py:23-31, shape (ndist, 2nz, 2n)."""
    fx = np.fft.fftfreq(2 * n, d=voxelsize).astype(dtype)
    fy = np.fft.fftfreq(2 * nz, d=voxelsize).astype(dtype)
    fx, fy = np.meshgrid(fx, fy)
    f2 = fx ** 2 + fy ** 2
    norm = float(4 * n * nz)
    out = np.empty((len(distance), 2 * nz, 2 * n), dtype=_cdt(dtype))
    for j, z in enumerate(distance):
        if dtype == np.float32:
            phase = np.float32(np.pi * wavelength * z) * f2       # weak-scalar FP32 product
            out[j] = (np.exp(-1j * phase.astype(np.complex64)) / np.float32(norm)).astype(np.complex64)
        else:
            out[j] = np.exp(-1j * np.pi * wavelength * z * f2) / norm
    return out


def fresnel_kernels_1d(n, nz, wavelength, voxelsize, distance, dtype=np.float32):
    """K_x (ndist, 2n) and K_y (ndist, 2nz) of processing/propagation_large.py:69-79."""
    fx = np.fft.fftfreq(2 * n, d=voxelsize).astype(dtype)
    fy = np.fft.fftfreq(2 * nz, d=voxelsize).astype(dtype)
    kx = np.empty((len(distance), 2 * n), dtype=_cdt(dtype))
    ky = np.empty((len(distance), 2 * nz), dtype=_cdt(dtype))
    for j, z in enumerate(distance):
        a = np.pi * wavelength * float(z)
        if dtype == np.float32:
            kx[j] = np.exp(-1j * (np.float32(a) * fx * fx).astype(np.complex64))
            ky[j] = np.exp(-1j * (np.float32(a) * fy * fy).astype(np.complex64))
        else:
            kx[j] = np.exp(-1j * a * fx * fx)
            ky[j] = np.exp(-1j * a * fy * fy)
    return kx, ky


def pad_symmetric(f):
    """``pad_fwd`` (processing/kernels.py:530-550): (b, nz, n) -> (b, 2nz, 2n)."""
    b, nz, n = f.shape
    tx = np.arange(2 * n)
    ty = np.arange(2 * nz)
    txx = np.where(tx < n // 2, n // 2 - tx - 1,
                   np.where(tx >= n + n // 2, 2 * n - tx + n // 2 - 1, tx - n // 2))
    tyy = np.where(ty < nz // 2, nz // 2 - ty - 1,
                   np.where(ty >= nz + nz // 2, 2 * nz - ty + nz // 2 - 1, ty - nz // 2))
    return f[:, tyy[:, None], txx[None, :]]


def pad_symmetric_T(g, nz, n):
    """Transpose of ``pad_symmetric``: fold-add (b, 2nz, 2n) -> (b, nz, n)."""
    b = g.shape[0]
    tx = np.arange(2 * n)
    ty = np.arange(2 * nz)
    txx = np.where(tx < n // 2, n // 2 - tx - 1,
                   np.where(tx >= n + n // 2, 2 * n - tx + n // 2 - 1, tx - n // 2))
    tyy = np.where(ty < nz // 2, nz // 2 - ty - 1,
                   np.where(ty >= nz + nz // 2, 2 * nz - ty + nz // 2 - 1, ty - nz // 2))
    out = np.zeros((b, nz, n), dtype=g.dtype)
    tmp = np.zeros((b, 2 * nz, n), dtype=g.dtype)
    for x in range(2 * n):
        tmp[:, :, txx[x]] += g[:, :, x]
    for y in range(2 * nz):
        out[:, tyy[y], :] += tmp[:, y, :]
    return out


def _as3(psi):
    psi = np.asarray(psi)
    return (psi[None], True) if psi.ndim == 2 else (psi, False)


def pad_zero(f):
    """Zero padding (b, nz, n) -> (b, 2nz, 2n), frame centred (the build's pad="zero" mode)."""
    b, nz, n = f.shape
    out = np.zeros((b, 2 * nz, 2 * n), dtype=f.dtype)
    out[:, nz // 2:nz // 2 + nz, n // 2:n // 2 + n] = f
    return out


def D(psi, fker_j, dtype=np.float32, pad="symmetric"):
    """``Propagation.D`` (processing/propagation.py:47-64) with one kernel fker[j]."""
    psi, added = _as3(psi)
    cdt = _cdt(dtype)
    _, nz, n = psi.shape
    big = (pad_symmetric if pad == "symmetric" else pad_zero)(psi.astype(cdt))
    big = sfft.fft2(big, axes=(-2, -1)).astype(cdt)
    big = (big * fker_j).astype(cdt)
    big = sfft.ifft2(big, axes=(-2, -1), norm="forward").astype(cdt)
    out = big[:, nz // 2:nz // 2 + nz, n // 2:n // 2 + n].copy()
    return out[0] if added else out


def D_separable(psi, kx_j, ky_j, dtype=np.float32):
    """``PropagationLarge.D`` arithmetic (processing/propagation_large.py:239-348)."""
    psi, added = _as3(psi)
    cdt = _cdt(dtype)
    b, nz, n = psi.shape
    h = n // 2
    x = np.concatenate([psi[:, :, :h][:, :, ::-1], psi, psi[:, :, -h:][:, :, ::-1]], axis=2).astype(cdt)
    x = sfft.fft(x, axis=-1).astype(cdt)
    x = (x * kx_j).astype(cdt)
    hz = nz // 2
    y = np.concatenate([x[:, :hz][:, ::-1], x, x[:, -hz:][:, ::-1]], axis=1)
    y = sfft.fft(y, axis=1).astype(cdt)
    y = (y * ky_j[:, None]).astype(cdt)
    y = sfft.ifft(y, axis=1).astype(cdt)
    x = y[:, hz:hz + nz, :]
    x = sfft.ifft(x, axis=-1).astype(cdt)
    out = x[:, :, h:h + n].copy()
    return out[0] if added else out


def DT(phi, fker_j, dtype=np.float32, pad="symmetric"):
    """Adjoint of ``D``: crop^T, FFT2, * conj(fker), IFFT2(norm="forward"), pad^T."""
    phi, added = _as3(phi)
    cdt = _cdt(dtype)
    b, nz, n = phi.shape
    big = np.zeros((b, 2 * nz, 2 * n), dtype=cdt)
    big[:, nz // 2:nz // 2 + nz, n // 2:n // 2 + n] = phi
    # the adjoint of the unnormalised inverse DFT is the unnormalised forward DFT
    # and vice versa; fker's real 1/(4 n nz) factor is kept by the conjugate
    big = sfft.fft2(big, axes=(-2, -1)).astype(cdt)
    big = (big * np.conj(fker_j)).astype(cdt)
    big = sfft.ifft2(big, axes=(-2, -1), norm="forward").astype(cdt)
    if pad == "symmetric":
        out = pad_symmetric_T(big, nz, n)
    else:
        out = big[:, nz // 2:nz // 2 + nz, n // 2:n // 2 + n].copy()
    return out[0] if added else out


def psi_from_proj(delta_raw, inv_norm, inv_beta_ratio, dtype=np.float32):
    """``_psi_from_proj`` (step3_propagation.py:94-103)."""
    p = np.asarray(delta_raw).astype(dtype)
    phase = p * dtype(inv_norm)
    atten = np.exp(-phase * dtype(inv_beta_ratio))
    return (atten * np.cos(phase) + 1j * (atten * np.sin(phase))).astype(_cdt(dtype))


def abs2(z):
    """``_abs2_c64_to_f32`` (step3_propagation.py:105-110)."""
    z = np.asarray(z)
    return (z.real * z.real + z.imag * z.imag).astype(z.real.dtype)
