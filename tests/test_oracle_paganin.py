"""Pins of the Paganin oracle against closed forms (no GPU)."""
import numpy as np
import pytest

from oracle import paganin as pg

WL, VOX, ALPHA = 4.13e-11, 1.4e-6, 1e-3


def test_constant_frame_and_dc_gain():
    n, nz = 32, 16
    x = np.full((2, nz, n), 0.73)
    assert np.allclose(pg.retrieve(x, WL, VOX, 0.5, ALPHA), -np.log(0.73), rtol=0, atol=1e-14)
    assert pg.filt(n, nz, WL, VOX, 0.5, ALPHA)[0, 0] == 1.0


def test_zero_distance_is_minus_log():
    x = np.random.default_rng(0).uniform(0.2, 1.5, (3, 16, 24))
    assert np.allclose(pg.retrieve(x, WL, VOX, 0.0, ALPHA), -np.log(x), rtol=0, atol=1e-13)


@pytest.mark.parametrize("p,q,dist", [(1, 0, 0.1), (0, 3, 1.0), (2, 5, 0.02), (-3, 7, 0.5), (8, 16, 1.0)])
def test_single_mode_closed_form(p, q, dist):
    """I = c + e cos(2 pi (p y / nz + q x / n)) -> -log(c + e H(p, q) cos(...)): fixes the filter's form, its
    sign, the 4 pi and the frequency scale 1 / (N d) on both axes (the Nyquist column q = n/2 included)."""
    n, nz, c, e = 32, 16, 1.0, 0.3
    y, x = np.mgrid[0:nz, 0:n]
    phase = 2 * np.pi * (p * y / nz + q * x / n)
    inten = c + e * np.cos(phase)
    h = pg.mode_gain(p, q, n, nz, WL, VOX, dist, ALPHA)
    assert 0 < h < 1
    assert np.allclose(pg.retrieve(inten, WL, VOX, dist, ALPHA), -np.log(c + e * h * np.cos(phase)), rtol=0,
                       atol=1e-12)


def test_clip_at_1e_minus_6():
    x = np.zeros((16, 16))
    assert np.allclose(pg.retrieve(x, WL, VOX, 0.3, ALPHA), -np.log(1e-6))


def test_fp32_restatement_close_to_fp64():
    x = np.random.default_rng(1).uniform(0.5, 1.0, (2, 32, 64))
    a = pg.retrieve(x, WL, VOX, 0.1, ALPHA, np.float32)
    b = pg.retrieve(x, WL, VOX, 0.1, ALPHA, np.float64)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5
