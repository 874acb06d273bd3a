"""Seeded synthetic inputs: ellipsoid phantoms, smooth random projection fields, angles.

Parameters are drawn with ``numpy.random.default_rng(seed)`` so that every run
sees the same volume; the volume itself is evaluated with torch on the
requested device (GPU for large sizes).
"""
from __future__ import annotations

import numpy as np
import torch


def angles(ntheta: int, span: float = np.pi) -> np.ndarray:
    """np.linspace(0, span, ntheta, endpoint=False) as float32 (step2_radon.py:106)."""
    return np.linspace(0.0, span, ntheta, endpoint=False).astype(np.float32)


def ellipsoid_params(nz: int, n: int, count: int = 12, seed: int = 0, vrange=(1e-7, 1e-6), semi=(0.05, 0.4)):
    """Centres uniform in the disc inscribed in the (n, n) plane, z centres in [0, nz),
    semi-axes uniform in semi*n (z semi-axis scaled to nz), values uniform in vrange."""
    rng = np.random.default_rng(seed)
    r = 0.5 * n * np.sqrt(rng.uniform(0, 1, count))
    phi = rng.uniform(0, 2 * np.pi, count)
    cx = n / 2 + r * np.cos(phi)
    cy = n / 2 + r * np.sin(phi)
    cz = rng.uniform(0, nz, count)
    ax = rng.uniform(semi[0], semi[1], count) * n
    ay = rng.uniform(semi[0], semi[1], count) * n
    az = rng.uniform(semi[0], semi[1], count) * nz
    val = rng.uniform(vrange[0], vrange[1], count)
    return [dict(c=(cz[i], cy[i], cx[i]), a=(az[i], ay[i], ax[i]), v=val[i]) for i in range(count)]


def ellipsoids(nz: int, n: int, count: int = 12, seed: int = 0, noise: float = 0.0, device=None,
               vrange=(1e-7, 1e-6), rows=None) -> torch.Tensor:
    """(nz, n, n) float32 sum of seeded ellipsoids; ``noise`` adds a Gaussian
    texture of that relative strength (of the peak value) on the support.
    ``rows=(r0, r1)`` evaluates only rows [r0, r1) of the nz-row phantom (a
    z-stack of a taller sample; neighbouring stacks may overlap)."""
    device = torch.device(device) if device is not None else torch.device("cpu")
    r0, r1 = (0, nz) if rows is None else rows
    vol = torch.zeros((r1 - r0, n, n), dtype=torch.float32, device=device)
    for e in ellipsoid_params(nz, n, count, seed, vrange):
        (cz, cy, cx), (az, ay, ax), v = e["c"], e["a"], e["v"]
        cz -= r0
        z0, z1 = max(0, int(cz - az)), min(r1 - r0, int(cz + az) + 1)
        y0, y1 = max(0, int(cy - ay)), min(n, int(cy + ay) + 1)
        x0, x1 = max(0, int(cx - ax)), min(n, int(cx + ax) + 1)
        if z0 >= z1 or y0 >= y1 or x0 >= x1:
            continue
        zz = ((torch.arange(z0, z1, device=device, dtype=torch.float32) + 0.5 - cz) / az) ** 2
        yy = ((torch.arange(y0, y1, device=device, dtype=torch.float32) + 0.5 - cy) / ay) ** 2
        xx = ((torch.arange(x0, x1, device=device, dtype=torch.float32) + 0.5 - cx) / ax) ** 2
        inside = (zz[:, None, None] + yy[None, :, None] + xx[None, None, :]) <= 1.0
        vol[z0:z1, y0:y1, x0:x1] += inside.to(torch.float32) * float(v)
    if noise > 0:
        g = torch.Generator(device=device).manual_seed(seed + 1)
        tex = torch.randn(vol.shape, generator=g, device=device, dtype=torch.float32)
        peak = float(vol.max()) if vol.numel() else 0.0
        vol += (vol > 0).to(torch.float32) * tex * (noise * peak)
    return vol


def smooth_field(nframes: int, nz: int, n: int, peak: float = 2.0, seed: int = 0, cutoff: float = 0.05,
                 device=None) -> torch.Tensor:
    """(nframes, nz, n) float32 low-pass-filtered Gaussian noise scaled to ``peak``."""
    device = torch.device(device) if device is not None else torch.device("cpu")
    g = torch.Generator(device=device).manual_seed(seed)
    x = torch.randn((nframes, nz, n), generator=g, device=device, dtype=torch.float32)
    fy = torch.fft.fftfreq(nz, device=device)[:, None]
    fx = torch.fft.rfftfreq(n, device=device)[None, :]
    filt = torch.exp(-(fx ** 2 + fy ** 2) / (2 * cutoff ** 2))
    out = torch.empty_like(x)
    for i in range(nframes):
        out[i] = torch.fft.irfft2(torch.fft.rfft2(x[i]) * filt, s=(nz, n))
    out -= out.amin()
    out *= peak / max(float(out.amax()), 1e-30)
    return out


def uneven_mosaic(seed: int = 7):
    """Uneven mosaic scan: 4 x-tiles x 9 z-stacks of 96 rows (8-row overlaps between stacks).  The x-tiles'
    detector widths are a seeded draw without replacement from {1536, 2048, 3072, 4096} (so the batch fits one
    GPU: ~113 GB of inputs, ~57 GB of outputs); each tile's angle count is drawn from {1029, 2058}.  Returns
    one dict per tile: x, s (stack), n, nz, ntheta, rows (row range in the full sample) and ztotal."""
    rng = np.random.default_rng(seed)
    nx, nzs, rows, overlap = 4, 9, 96, 8
    widths = rng.permutation([1536, 2048, 3072, 4096])[:nx]
    tiles = []
    for i in range(nx):
        for s in range(nzs):
            nt = int(rng.choice([1029, 2058]))
            tiles.append(dict(x=i, s=s, n=int(widths[i]), nz=rows, ntheta=nt,
                              rows=(s * (rows - overlap), s * (rows - overlap) + rows),
                              ztotal=nzs * (rows - overlap) + overlap))
    return tiles
