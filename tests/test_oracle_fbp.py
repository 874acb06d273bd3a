"""Pins of the FBP-filter oracle (no GPU)."""
import numpy as np
import pytest

from oracle import fbp


@pytest.mark.parametrize("N", [25, 33, 40])
def test_wint_moments_closed_form(N):
    """Without the taper the weights integrate t * t^d over [0, T] exactly up to d = 11 (to 1e-8: the
    inverse Vandermonde on nodes from 1e-40 costs ~1e-9; a wrong window, weight or overlap factor is O(1))."""
    t = np.arange(N) / (4.0 * (N - 1))
    w = fbp.wint(12, t)
    T = t[-1]
    for d in range(0, 12):
        assert abs(np.sum(w * t ** d) - T ** (d + 2) / (d + 2)) < 1e-8 * T ** (d + 2) / (d + 2), d


def test_wint_taper_is_a_line_through_origin():
    t = np.arange(0, 257) / 512.0
    w = fbp.wint(12, t)
    N = len(t)
    slope = w[-40] / (N - 40)
    assert np.allclose(w[-40:], slope * np.arange(N - 40, N), rtol=0, atol=1e-15)


@pytest.mark.parametrize("n", [16, 64])
def test_none_filter_closed_form(n):
    """'none' weights: 3 n ne everywhere but 12 n ne at DC -> 3 n ne x + 9 n sum(edge_pad(x))."""
    ne = 4 * n
    w = fbp.calc_filter(n, "none")
    assert np.all(w[1:] == np.float32(3 * n * ne)) and w[0] == np.float32(12 * n * ne)
    x = np.random.default_rng(n).standard_normal((3, n))
    got = fbp.filter(x, w.astype(np.float64))
    ref = 3 * n * ne * x + 9 * n * fbp.edge_pad(x, ne).sum(-1, keepdims=True)
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())


def test_filter_is_circular_convolution_of_padded_row():
    """Brute force for a small n (the quadrature needs ne / 2 + 1 >= 23 points): irfft(rfft(p) w) = p
    circularly convolved with the real even kernel ifft(w_even)."""
    n = 16
    ne = 4 * n
    w = fbp.calc_filter(n, "shepp").astype(np.float64)
    weven = np.concatenate([w, w[1:-1][::-1]])
    kern = np.real(np.fft.ifft(weven))
    x = np.random.default_rng(1).standard_normal(n)
    p = fbp.edge_pad(x, ne)
    conv = np.array([sum(p[j] * kern[(i - j) % ne] for j in range(ne)) for i in range(ne)])
    pad = (ne - n) // 2
    assert np.allclose(fbp.filter(x, w), conv[pad:pad + n], rtol=1e-10, atol=1e-10)


def test_edge_pad_matches_numpy_edge_mode():
    """edge_pad (the reference's edge replication to ne = 4 n) = np.pad(mode="edge") with (ne - n) // 2 on the
    left and the rest on the right."""
    x = np.random.default_rng(3).standard_normal((2, 3, 24))
    for ne in (96, 97):
        pad = (ne - 24) // 2
        ref = np.pad(x, ((0, 0), (0, 0), (pad, ne - 24 - pad)), mode="edge")
        assert np.array_equal(fbp.edge_pad(x, ne), ref)
