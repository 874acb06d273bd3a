"""FBPFilter: the 1-D FBP filter of processing/fbp_filter.py:39-242 on one MI355X.

The sinogram rows (last axis, length n) are edge-replicated to ne = 4 n,
transformed, multiplied by real weights w (ne // 2 + 1 values, calc_filter),
transformed back (1 / ne) and cropped to n, in place
(processing/fbp_filter.py:128-210).  On the GPU two rows ride in one complex
FFT of length ne (the weights are real and even, so the two real results come
back as the real and imaginary parts), with the edge padding, the weights and
the crop fused into one kernel (mm_fbp_filter): one read and one write of the
sinogram.  The weights are built on the host (calc_filter_host): the ramp
|t| is integrated by a 12-point piecewise-Lagrange rule (ramp_quadrature, the
rule processing/fbp_filter.py:94-126 defines) and shaped by the named window.

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


def _window_moments(order: int):
    """Integrals of the Lagrange basis on `order` equispaced nodes s_i = i / (order - 1) of [0, 1].

    Returns (M0, M1), each (order, order - 1): M0[i, k] = int L_i(s) ds and M1[i, k] = int s L_i(s) ds over
    the k-th gap [s_k, s_{k+1}].  L_i is evaluated in product form with barycentric weights and integrated
    by Gauss-Legendre on each gap with enough points to be exact for its degree (order)."""
    s_nodes = np.arange(order) / (order - 1)
    gaps = s_nodes[:, None] - s_nodes[None, :]
    np.fill_diagonal(gaps, 1.0)
    bary = 1.0 / gaps.prod(axis=1)
    gx, gw = np.polynomial.legendre.leggauss(order // 2 + 2)
    M0 = np.empty((order, order - 1))
    M1 = np.empty((order, order - 1))
    for k in range(order - 1):
        lo, hi = s_nodes[k], s_nodes[k + 1]
        s = lo + (hi - lo) * 0.5 * (gx + 1.0)
        ds = (hi - lo) * 0.5 * gw
        basis = np.stack([bary[i] * np.prod(s[None, :] - np.delete(s_nodes, i)[:, None], axis=0)
                          for i in range(order)])
        M0[:, k] = basis @ ds
        M1[:, k] = basis @ (ds * s)
    return M0, M1


def ramp_quadrature(t: np.ndarray, order: int = 12, taper: int = 40) -> np.ndarray:
    """Weights q with sum_i q_i f(t_i) ~ int_{t_0}^{t_end} t f(t) dt on the increasing grid t: the integrand
    of the ramp filter |t| (processing/fbp_filter.py:94-126 defines the rule).

    Every run of `order` consecutive grid points is a window; on each window f is replaced by its
    Lagrange interpolant, and the integral of t times that interpolant over each gap of the window is
    exact (window moments M0, M1 scaled to the window: t = t_j + h s, h = t_{j+order-1} - t_j).  A gap
    covered by c windows takes the mean of their c values.  The last `taper` weights are then replaced by
    the line through the origin that meets the weight at index N - taper."""
    t = np.asarray(t, dtype=np.float64)
    N = len(t)
    nwin = N - order + 1
    if nwin < 1:
        raise ValueError(f"ramp_quadrature needs at least {order} grid points, got {N}")
    M0, M1 = _window_moments(order)
    gap = np.arange(N - 1)
    covering = np.minimum(gap, nwin - 1) - np.maximum(0, gap - (order - 2)) + 1
    share = np.lib.stride_tricks.sliding_window_view(1.0 / covering, order - 1)   # (nwin, order - 1)
    j = np.arange(nwin)
    h = t[j + order - 1] - t[j]
    per_window = (h * h)[:, None] * (share @ M1.T) + (h * t[j])[:, None] * (share @ M0.T)   # (nwin, order)
    q = np.zeros(N)
    for i in range(order):
        q[i:i + nwin] += per_window[:, i]
    if N > taper:
        q[N - taper:] = q[N - taper] * np.arange(N - taper, N) / (N - taper)
    return q


# Apodisation windows of the named filters as functions of x = t / d (d = 0.5, the Nyquist frequency of the
# unpadded row), each with its support (processing/fbp_filter.py:48-91 names the filters and their shapes).
_WINDOWS = {
    "ramp": (lambda x: np.ones_like(x), np.inf),
    "shepp": (lambda x: np.sinc(x / 2.0), 2.0),
    "cosine": (lambda x: np.cos(0.5 * np.pi * x), 1.0),
    "cosine2": (lambda x: np.cos(0.5 * np.pi * x) ** 2, 1.0),
    "hamming": (lambda x: 0.54 + 0.46 * np.cos(np.pi * x), 1.0),
    "hann": (lambda x: 0.5 * (1.0 + np.cos(np.pi * x)), 1.0),
    "parzen": (lambda x: (1.0 - x) ** 3, 1.0),
}


def calc_filter_host(n: int, name: str) -> np.ndarray:
    """(4n // 2 + 1,) float32 half-spectrum weights of the named filter on the 4n-padded grid.

    Frequencies t_k = k / (4n), k = 0 .. 2n.  The ramp part is (4n / 2) * ramp_quadrature(t), 'none'
    replaces it by the constant 4n / 2 with a doubled DC value; the named window multiplies it on its
    support.  Negative values are clipped, everything is doubled, DC doubled again, and the result is
    scaled by 3 n (processing/fbp_filter.py:48-91 fixes these factors)."""
    ne = 4 * n
    t = np.arange(ne // 2 + 1) / ne
    half = 0.5 * ne
    if name == "none":
        resp = np.full(t.shape, half)
        resp[0] *= 2.0
    elif name in _WINDOWS:
        shape, support = _WINDOWS[name]
        x = t / 0.5
        resp = half * ramp_quadrature(t) * np.where(x <= support, shape(np.minimum(x, support)), 0.0)
    else:
        raise ValueError(f"FBPFilter: no filter named {name!r} (known: {', '.join(FILTERS)})")
    resp = 2.0 * np.maximum(resp, 0.0)
    resp[0] *= 2.0
    return (3 * n * resp).astype(np.float32)


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
        self._tables: list[tuple] = []   # (tensor or None, version, weight bytes, (G, A, B))
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
        self.check_weights(w)
        rows = sino.numel() // self.n
        if self.method == "4n":
            wd = as_device_tensor(w, torch.float32, sino.device).contiguous()
            _lib.check(self._lib.mm_fbp_filter(sino.data_ptr(), wd.data_ptr(), rows, self.n, _lib.stream_handle()),
                       "FBPFilter.filter")
            return sino
        G, A, B = self._two_n(w)
        _lib.check(self._lib.mm_fbp_filter_2n(sino.data_ptr(), G.data_ptr(), A.data_ptr(), B.data_ptr(), rows,
                                              self.n, _lib.stream_handle()), "FBPFilter.filter")
        return sino

    def check_weights(self, w):
        size = w.numel() if isinstance(w, torch.Tensor) else np.size(w)
        if size != self.ne // 2 + 1:
            raise ValueError(f"filter weights have {size} values, expected {self.ne // 2 + 1}")

    _TABLES_KEPT = 4

    def _two_n(self, w):
        """Device (G, A, B) of the length-2n form for the weights w (numpy array or tensor).

        Tables are keyed by the weight values (their float32 bytes), never by a tensor's address, so a new
        tensor at a reused address, or w changed in place, gets its own tables.  A tensor seen before (the
        entry keeps a reference to it, so its address cannot be reused while cached) at the same version
        counter skips the host copy.  The last few weight sets are kept."""
        if isinstance(w, torch.Tensor):
            for i, (ref, ver, _, tables) in enumerate(self._tables):
                if ref is w and ver == w._version:
                    return self._promote(i)
            host = w.detach().to("cpu", torch.float32).contiguous().numpy()
            ref, ver = w, w._version
        else:
            host = np.ascontiguousarray(np.asarray(w, dtype=np.float32))
            ref, ver = None, -1
        key = host.tobytes()
        for i, (_, _, k, tables) in enumerate(self._tables):
            if k == key:
                self._tables[i] = (ref, ver, k, tables)
                return self._promote(i)
        G, A, B = two_n_tables(host, self.n)
        tables = tuple(torch.from_numpy(a).to(self.device) for a in (G, A, B))
        self._tables.insert(0, (ref, ver, key, tables))
        del self._tables[self._TABLES_KEPT:]
        return tables

    def _promote(self, i):
        if i:
            self._tables.insert(0, self._tables.pop(i))
        return self._tables[0][3]

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
