"""Pins the NumPy USFFT oracle (oracle/usfft.py) against mathematics it does not share code with."""
import numpy as np
import pytest

from oracle import independent as ind
from oracle import usfft


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


N, NZ, NT = 64, 2, 24
THETA = np.linspace(0, np.pi, NT, endpoint=False).astype(np.float32)
OBJ = np.random.default_rng(0).standard_normal((NZ, N, N)).astype(np.float32)


@pytest.fixture(scope="module")
def r64():
    return usfft.tomo_R(OBJ, THETA, np.float64)


def test_params_n_independent():
    # processing/tomo.py:39-42: m = 4 for every n in the build's range
    for n in (16, 64, 256, 1024, 2048, 4096, 8192):
        eps, mu, m = usfft.usfft_params(n)
        assert m == 4 and eps == 1e-3
        assert np.isclose(mu * 2 * n * n, -np.log(1e-3))


def test_R_vs_exact_nudft(r64):
    # the m=4 Gaussian window approximates the exact polar NUDFT to O(eps)
    err = rel(r64, ind.radon_from_nudft(OBJ, THETA))
    assert err < 1.5e-3, err


def test_R_theta0_is_column_sum(r64):
    err = rel(r64[0], OBJ.sum(axis=1) / np.sqrt(N * NT))
    assert err < 1.5e-3, err


def test_R_theta_half_pi_is_reversed_row_sum():
    th = np.array([np.pi / 2], dtype=np.float32)
    r = usfft.tomo_R(OBJ, th, np.float64)
    rs = OBJ.sum(axis=2) / np.sqrt(N * 1)
    idx = (N - np.arange(N)) % N
    assert rel(r[0], rs[:, idx]) < 1.5e-3
    assert rel(r[0], rs) > 0.5   # a transposed or unreversed convention fails


def test_R_gaussian_blob_line_integrals():
    cent = [(5.0, -7.0), (-10.0, 3.0)]
    sig = [2.5, 3.0]
    amp = [1.0, 0.5]
    th = np.linspace(0, 2 * np.pi, 17, endpoint=False).astype(np.float32)
    ob = ind.gaussian_blob(N, 1, cent, sig, amp)
    r = usfft.tomo_R(ob, th, np.float64)[:, 0]
    p = ind.gaussian_blob_projection(N, th, cent, sig, amp) / np.sqrt(N * len(th))
    assert rel(r, p) < 1e-4
    # the opposite y0 sign convention would mirror the projections
    p_wrong = ind.gaussian_blob_projection(N, -th, cent, sig, amp) / np.sqrt(N * len(th))
    assert rel(r, p_wrong) > 0.1


@pytest.mark.parametrize("span", [np.pi, 2 * np.pi])
def test_adjoint_identity_real_fp64(span):
    th = np.linspace(0, span, NT, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((NZ, N, N))
    y = rng.standard_normal((NT, NZ, N))
    lhs = np.sum(usfft.tomo_R(x, th, np.float64) * y)
    rhs = np.sum(x * usfft.tomo_RT(y, th, dtype=np.float64))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_complex_RT_linear_and_real_part():
    # RT of a complex64 sinogram is complex-linear and its real-input real part is the float32 path
    rng = np.random.default_rng(2)
    yr = rng.standard_normal((NT, NZ, N))
    yi = rng.standard_normal((NT, NZ, N))
    full = usfft.tomo_RT(yr + 1j * yi, THETA, dtype=np.float64)
    a = usfft.tomo_RT(yr.astype(np.complex128), THETA, dtype=np.float64)
    b = usfft.tomo_RT(yi.astype(np.complex128), THETA, dtype=np.float64)
    assert rel(full, a + 1j * b) < 1e-12
    assert rel(a.real, usfft.tomo_RT(yr, THETA, dtype=np.float64)) < 1e-12
    assert np.linalg.norm(a.imag) > 1e-6 * np.linalg.norm(a.real)   # the imaginary part is not dropped


def test_fp32_restatement_close_to_fp64(r64):
    r32 = usfft.tomo_R(OBJ, THETA, np.float32)
    assert r32.dtype == np.float32
    assert rel(r32, r64) < 1e-5
    y = np.random.default_rng(3).standard_normal((NT, NZ, N)).astype(np.float32)
    assert rel(usfft.tomo_RT(y, THETA, dtype=np.float32), usfft.tomo_RT(y, THETA, dtype=np.float64)) < 1e-5


def test_large_restatement_same_operator(r64):
    # TomoLargeReal's separable phi / scale placement / clamped compact gather: same operator to ~1e-7
    assert rel(usfft.tomo_large_R(OBJ, THETA, np.float64), r64) < 1e-6
    yc = np.random.default_rng(4).standard_normal((NT, NZ, N)) + 1j
    assert rel(usfft.tomo_large_RT(yc, THETA, np.float64), usfft.tomo_RT(yc, THETA, dtype=np.float64)) < 1e-12


def test_planted_mistakes_fail():
    """A dropped centring sign or a missing conj in the reflection is caught by the NUDFT pin."""
    nud = ind.radon_from_nudft(OBJ[:1], THETA)
    good = usfft.tomo_R(OBJ[:1], THETA, np.float64)
    assert rel(good, nud) < 1.5e-3
    orig = usfft.gather_half

    def no_conj(fde, theta, n, dtype=np.float32, variant="real"):
        fde2 = np.asarray(fde).copy()
        return orig(np.conj(fde2), theta, n, dtype, variant)

    usfft.gather_half = no_conj
    try:
        bad = usfft.tomo_R(OBJ[:1], THETA, np.float64)
    finally:
        usfft.gather_half = orig
    assert rel(bad, nud) > 0.1
