"""FBPFilter: the 1-D FBP filter of processing/fbp_filter.py:39-242 on one MI355X.

The sinogram rows (last axis, length n) are edge-replicated to ne = 4 n,
transformed, multiplied by real weights w (ne // 2 + 1 values, calc_filter),
transformed back (1 / ne) and cropped to n, in place
(processing/fbp_filter.py:128-210).  On the GPU two rows ride in one complex
FFT of length ne (the weights are real and even, so the two real results come
back as the real and imaginary parts), with the edge padding, the weights and
the crop fused into one kernel (mm_fbp_filter): one read and one write of the
sinogram.  The weights are computed on the host exactly as the reference does
(calc_filter with the 12-point quadrature _wint, :48-126).

By default the filter runs on transforms of length 2 n instead of 4 n
(mm_fbp_filter_2n): with h = irfft(w, 4 n) the edge-padded convolution is a
circular convolution of length 2 n with h cut to the lags |l| < n, plus the
two edge values times fixed vectors A, B (the sums of h over the padding).
``FBPFilter(n, method="4n")`` keeps the padded transforms of the reference.
"""
from __future__ import annotations

import numpy as np
import torch

from model_mosaic_amd import _lib
from model_mosaic_amd._carriers import as_device_tensor, require_gpu

FILTERS = ("none", "ramp", "shepp", "cosine", "cosine2", "hamming", "hann", "parzen")


def wint(order: int, t: np.ndarray) -> np.ndarray:
    """Quadrature weights of the ramp |t| on the grid t (processing/fbp_filter.py:94-126): for each window
    of `order` consecutive grid points, the integrals of x * p_k(x) and p_k(x) over the short intervals of
    the window, p_k the Lagrange basis, averaged over the windows that cover a point; the last 40 values
    are replaced by a linear taper."""
    N = len(t)
    s = np.linspace(1e-40, 1, order)
    k1 = np.arange(order)
    k2 = np.arange(1, order + 2)
    iv = np.linalg.inv(np.exp(np.outer(k1, np.log(s))))
    u = np.diff(np.exp(np.outer(k2, np.log(s))) * np.tile(1.0 / k2[:, None], [1, order]))
    W1 = iv @ u[1:order + 1, :]
    W2 = iv @ u[0:order, :]
    p = 1.0 / np.concatenate((np.arange(1, order), (order - 1) * np.ones(N - 2 * (order - 1) - 1), np.arange(order - 1, 0, -1)))
    w = np.zeros(N)
    for j in range(N - order + 1):
        h = t[j + order - 1] - t[j]
        W = h * h * W1 + h * t[j] * W2
        for k in range(order - 1):
            w[j:j + order] += p[j + k] * W[:, k]
    if N > 40:
        w[-40:] = w[-40] / (N - 40) * np.arange(N - 40, N)
    return w


def two_n_tables(w: np.ndarray, n: int):
    """(G, A, B) float32 of the length-2n form of the filter with half-spectrum weights w (4n // 2 + 1):
    h = irfft(w, 4n) is real and even; G = rfft(g) / (2n) (real, n + 1 values) for g = h on the lags
    |l| < n laid out mod 2n; A[i] = sum_{d = i+1}^{pad+i} h[d] (left padding, pad = 3n/2 values of row[0])
    and B[i] = A[n - 1 - i] (right padding, row[n-1]).  Computed in FP64."""
    ne, L = 4 * n, 2 * n
    pad = (ne - n) // 2
    h = np.fft.irfft(np.asarray(w, dtype=np.float64), ne)
    g = np.zeros(L)
    lags = np.arange(-(n - 1), n)
    g[lags % L] = h[lags % ne]
    G = np.fft.rfft(g).real / L
    cs = np.cumsum(h)
    i = np.arange(n)
    A = cs[pad + i] - cs[i]
    B = A[::-1].copy()
    return G.astype(np.float32), A.astype(np.float32), B.astype(np.float32)


def calc_filter_host(n: int, name: str) -> np.ndarray:
    """(4n // 2 + 1,) float32 weights of the named filter (processing/fbp_filter.py:48-91)."""
    d = 0.5
    ne = 4 * n
    t = np.arange(0, ne // 2 + 1) / ne
    if name == "none":
        wfa = ne * 0.5 + t * 0
        wfa[0] *= 2
    elif name == "ramp":
        wfa = ne * 0.5 * wint(12, t)
    elif name == "shepp":
        wfa = ne * 0.5 * wint(12, t) * np.sinc(t / (2 * d)) * (t / d <= 2)
    elif name == "cosine":
        wfa = ne * 0.5 * wint(12, t) * np.cos(np.pi * t / (2 * d)) * (t / d <= 1)
    elif name == "cosine2":
        wfa = ne * 0.5 * wint(12, t) * np.cos(np.pi * t / (2 * d)) ** 2 * (t / d <= 1)
    elif name == "hamming":
        wfa = ne * 0.5 * wint(12, t) * (0.54 + 0.46 * np.cos(np.pi * t / d)) * (t / d <= 1)
    elif name == "hann":
        wfa = ne * 0.5 * wint(12, t) * (1 + np.cos(np.pi * t / d)) / 2.0 * (t / d <= 1)
    elif name == "parzen":
        wfa = ne * 0.5 * wint(12, t) * (1 - t / d) ** 3 * (t / d <= 1)
    else:
        raise ValueError(f"unknown FBP filter '{name}'; expected one of {', '.join(FILTERS)}")
    wfa = 2 * wfa * (wfa >= 0)
    wfa[0] *= 2
    return (3 * n * wfa).astype(np.float32)


class FBPFilter:
    """Same interface as processing/fbp_filter.py:39 (``FBPFilter(n, batch_chunk)``, ``calc_filter``,
    ``filter``, ``filter_host``); carriers are torch ROCm tensors."""

    def __init__(self, n: int, batch_chunk: int = 256, method: str = "2n"):
        require_gpu()
        if method not in ("2n", "4n"):
            raise ValueError(f"method must be '2n' or '4n', got {method!r}")
        self._lib = _lib.load()
        self.n = int(n)
        self.ne = 4 * self.n
        self.pad = (self.ne - self.n) // 2
        self.batch_chunk = int(batch_chunk)
        self.method = method
        self._cache: dict[str, torch.Tensor] = {}
        self._tables: dict[tuple, tuple] = {}
        self.device = torch.device("cuda", torch.cuda.current_device())

    def calc_filter(self, name: str) -> torch.Tensor:
        if name not in self._cache:
            self._cache[name] = torch.from_numpy(calc_filter_host(self.n, name)).to(self.device)
        return self._cache[name]

    def filter(self, sino: torch.Tensor, w: torch.Tensor, batch_chunk: int | None = None) -> torch.Tensor:
        """Filters ``sino`` (..., n) float32 in place along its last axis and returns it.  The whole
        batch is one launch (``batch_chunk`` is accepted for interface compatibility)."""
        if sino.shape[-1] != self.n:
            raise ValueError(f"sino last-axis {sino.shape[-1]} != filter n={self.n}")
        if not sino.is_contiguous():
            raise ValueError("sino must be C-contiguous (reshape view for chunk loop)")
        if sino.dtype != torch.float32 or not sino.is_cuda:
            raise ValueError("sino must be a float32 device tensor")
        w = as_device_tensor(w, torch.float32, sino.device).contiguous()
        if w.numel() != self.ne // 2 + 1:
            raise ValueError(f"filter weights have {w.numel()} values, expected {self.ne // 2 + 1}")
        rows = sino.numel() // self.n
        if self.method == "4n":
            _lib.check(self._lib.mm_fbp_filter(sino.data_ptr(), w.data_ptr(), rows, self.n, _lib.stream_handle()),
                       "FBPFilter.filter")
            return sino
        G, A, B = self._two_n(w)
        _lib.check(self._lib.mm_fbp_filter_2n(sino.data_ptr(), G.data_ptr(), A.data_ptr(), B.data_ptr(), rows,
                                              self.n, _lib.stream_handle()), "FBPFilter.filter")
        return sino

    def _two_n(self, w: torch.Tensor):
        """Device (G, A, B) of the length-2n form for weights w, built once per weight tensor (keyed by its
        storage and version counter, so an in-place change of w rebuilds them)."""
        key = (w.data_ptr(), w._version, w.numel())
        if key not in self._tables:
            G, A, B = two_n_tables(w.detach().cpu().numpy(), self.n)
            self._tables[key] = tuple(torch.from_numpy(a).to(w.device) for a in (G, A, B))
        return self._tables[key]

    def filter_host(self, sino_h: np.ndarray, w: torch.Tensor, batch_chunk: int | None = None) -> np.ndarray:
        """Host sinogram filtered in place, in slabs of ``batch_chunk`` rows (processing/fbp_filter.py:212)."""
        if sino_h.shape[-1] != self.n:
            raise ValueError(f"sino_h last-axis {sino_h.shape[-1]} != filter n={self.n}")
        if not sino_h.flags.c_contiguous:
            raise ValueError("sino_h must be C-contiguous (reshape view for chunk loop)")
        flat = sino_h.reshape(-1, self.n)
        bc = min(int(batch_chunk or self.batch_chunk), len(flat)) or 1
        for i0 in range(0, len(flat), bc):
            d = torch.from_numpy(np.ascontiguousarray(flat[i0:i0 + bc])).to(self.device)
            self.filter(d, w)
            flat[i0:i0 + bc] = d.cpu().numpy()
        return sino_h
