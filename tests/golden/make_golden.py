"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

The reference holds no stored arrays (its tests are relational only,
tests/test_large_variants.py), so these vectors come from the oracle of
oracle/, itself pinned by tests/test_oracle_*.py against closed forms, exact
NUDFT brute force and adjoint identities.  Regenerate with
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fresnel, usfft  # noqa: E402
from oracle.independent import radon_from_nudft  # noqa: E402


def phantom(nz, n, seed):
    """Seeded ellipsoid slices (same recipe as model_mosaic_amd.phantom, numpy only)."""
    rng = np.random.default_rng(seed)
    vol = np.zeros((nz, n, n), np.float64)
    yy, xx = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    for _ in range(12):
        r = 0.5 * n * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        cx, cy = n / 2 + r * np.cos(ph), n / 2 + r * np.sin(ph)
        ax, ay = rng.uniform(0.05, 0.4, 2) * n
        v = rng.uniform(1e-7, 1e-6)
        z = rng.integers(0, nz)
        vol[z] += v * (((xx - cx) / ax) ** 2 + ((yy - cy) / ay) ** 2 <= 1)
    return vol.astype(np.float32)


def radon_small():
    """n=64, nz=4, ntheta=24 on [0, 2 pi) (the reference test's geometry, tests/test_large_variants.py:52-54)."""
    n, nz, ntheta = 64, 4, 24
    theta = np.linspace(0, 2 * np.pi, ntheta, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(0)
    obj = rng.standard_normal((nz, n, n)).astype(np.float32)
    y = rng.standard_normal((ntheta, nz, n)).astype(np.float32)
    yc = (rng.standard_normal((ntheta, nz, n)) + 1j * rng.standard_normal((ntheta, nz, n))).astype(np.complex64)
    return dict(theta=theta, obj=obj, R=usfft.tomo_R(obj, theta, np.float64).astype(np.float32),
                y=y, RT=usfft.tomo_RT(y, theta, dtype=np.float64).astype(np.float32),
                yc=yc, RTc=usfft.tomo_RT(yc, theta, dtype=np.float64).astype(np.complex64))


def radon_c1():
    """CPU configuration sizes n=256, ntheta=384 on [0, pi), 2 ellipsoid slices."""
    n, nz, ntheta = 256, 2, 384
    theta = np.linspace(0, np.pi, ntheta, endpoint=False).astype(np.float32)
    obj = phantom(nz, n, 7)
    R64 = usfft.tomo_R(obj, theta, np.float64)
    return dict(theta=theta, obj=obj, R=R64.astype(np.float32),
                R_nudft=radon_from_nudft(obj[:1], theta).astype(np.float32),
                RT=usfft.tomo_RT(R64.astype(np.float32), theta, dtype=np.float64).astype(np.float32))


def fresnel_small():
    """n=64, nz=32, 3 frames, z in {0, 0.02, 0.1} m (tests/test_large_variants.py:180-188)."""
    n, nz, nt = 64, 32, 3
    wl, vox = 4.13e-11, 1.4e-6
    dist = np.array([0.0, 0.02, 0.1])
    rng = np.random.default_rng(0)
    psi = (rng.standard_normal((nt, nz, n)) + 1j * rng.standard_normal((nt, nz, n))).astype(np.complex64)
    fk = fresnel.fresnel_kernel_2d(n, nz, wl, vox, dist, np.float64)
    D = np.stack([fresnel.D(psi, fk[j], np.float64) for j in range(3)]).astype(np.complex64)
    DT = np.stack([fresnel.DT(psi, fk[j], np.float64) for j in range(3)]).astype(np.complex64)
    return dict(psi=psi, dist=dist, wavelength=np.float64(wl), voxelsize=np.float64(vox), D=D, DT=DT)


if __name__ == "__main__":
    for name, fn in [("radon_small", radon_small), ("radon_c1", radon_c1), ("fresnel_small", fresnel_small)]:
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print("wrote", path, os.path.getsize(path))
