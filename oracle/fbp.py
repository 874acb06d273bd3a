"""NumPy restatement of the FBP filter (TEST INFRASTRUCTURE ONLY).

Restates processing/fbp_filter.py:
  * ``calc_filter`` (:48-91) and the ramp quadrature ``_wint`` (:94-126);
  * ``FBPFilter.filter`` (:128-210): edge replication to ne = 4 n, rfft,
    multiplication by the weights, irfft (1 / ne), crop to n.

Parity pins (tests/test_oracle_fbp.py):
  * ``wint`` against moments computed in closed form: for grids short enough
    to skip the taper (N <= 40) the weights integrate t * t^d over [0, T]
    exactly for d <= 11 (sum_i w_i t_i^d = T^(d+2) / (d+2)); the taper of the
    last 40 values is a line through the origin (closed form);
  * ``filter`` with the 'none' weights against its closed form
    3 n ne x + 9 n sum(edge_pad(x)) (a scaled identity plus the doubled DC
    weight), and against a direct DFT convolution of the padded row for tiny n.
"""
from __future__ import annotations

import numpy as np


def wint(order, t):
    """processing/fbp_filter.py:94-126 in FP64."""
    N = len(t)
    s = np.linspace(1e-40, 1, order)
    k1 = np.arange(order)
    k2 = np.arange(1, order + 2)
    iv = np.linalg.inv(np.exp(np.outer(k1, np.log(s))))
    u = np.diff(np.exp(np.outer(k2, np.log(s))) * np.tile(1.0 / k2[:, None], [1, order]))
    W1 = iv @ u[1:order + 1, :]
    W2 = iv @ u[0:order, :]
    p = 1.0 / np.concatenate((np.arange(1, order), (order - 1) * np.ones(N - 2 * (order - 1) - 1),
                              np.arange(order - 1, 0, -1)))
    w = np.zeros(N)
    for j in range(N - order + 1):
        h = t[j + order - 1] - t[j]
        W = h * h * W1 + h * t[j] * W2
        for k in range(order - 1):
            w[j:j + order] += p[j + k] * W[:, k]
    if N > 40:
        w[-40:] = w[-40] / (N - 40) * np.arange(N - 40, N)
    return w


def calc_filter(n, name):
    """processing/fbp_filter.py:48-91: (4n // 2 + 1,) weights, float32."""
    d = 0.5
    ne = 4 * n
    t = np.arange(0, ne // 2 + 1) / ne
    ramp = lambda: ne * 0.5 * wint(12, t)
    shapes = {
        "none": lambda: ne * 0.5 + t * 0,
        "ramp": ramp,
        "shepp": lambda: ramp() * np.sinc(t / (2 * d)) * (t / d <= 2),
        "cosine": lambda: ramp() * np.cos(np.pi * t / (2 * d)) * (t / d <= 1),
        "cosine2": lambda: ramp() * np.cos(np.pi * t / (2 * d)) ** 2 * (t / d <= 1),
        "hamming": lambda: ramp() * (0.54 + 0.46 * np.cos(np.pi * t / d)) * (t / d <= 1),
        "hann": lambda: ramp() * (1 + np.cos(np.pi * t / d)) / 2.0 * (t / d <= 1),
        "parzen": lambda: ramp() * (1 - t / d) ** 3 * (t / d <= 1),
    }
    wfa = np.array(shapes[name](), dtype=np.float64)
    if name == "none":
        wfa[0] *= 2
    wfa = 2 * wfa * (wfa >= 0)
    wfa[0] *= 2
    return (3 * n * wfa).astype(np.float32)


def edge_pad(sino, ne):
    n = sino.shape[-1]
    pad = (ne - n) // 2
    return np.concatenate([np.repeat(sino[..., :1], pad, -1), sino, np.repeat(sino[..., -1:], ne - n - pad, -1)], -1)


def filter(sino, w, dtype=np.float64):
    """processing/fbp_filter.py:128-210 (out of place): edge pad, rfft, * w, irfft, crop."""
    sino = np.asarray(sino, dtype=dtype)
    n = sino.shape[-1]
    ne = 4 * n
    pad = (ne - n) // 2
    t = edge_pad(sino, ne)
    spec = np.fft.rfft(t, axis=-1) * np.asarray(w, dtype=dtype)
    out = np.fft.irfft(spec, n=ne, axis=-1)
    return out[..., pad:pad + n].astype(dtype)
