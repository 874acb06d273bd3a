"""Pins the NumPy Fresnel oracle (oracle/fresnel.py) against closed forms and identities."""
import numpy as np
import pytest

from oracle import fresnel as fr
from oracle import independent as ind


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


N, NZ, B = 64, 32, 3
WL, VOX = 4.13e-11, 1.4e-6
DIST = [0.0, 0.02, 0.1]
PSI = (np.random.default_rng(0).standard_normal((B, NZ, N))
       + 1j * np.random.default_rng(1).standard_normal((B, NZ, N)))


def test_pad_matches_numpy_symmetric():
    p = fr.pad_symmetric(PSI)
    q = np.pad(PSI, ((0, 0), (NZ // 2, NZ // 2), (N // 2, N // 2)), mode="symmetric")
    assert np.array_equal(p, q)


def test_zero_pad_matches_numpy_constant_and_z0_identity():
    """pad_zero = np.pad(mode="constant") with the frame centred, and D with zero padding at z = 0 is the
    identity (the chirp is 1, crop(pad(x)) = x)."""
    q = np.pad(PSI, ((0, 0), (NZ // 2, NZ // 2), (N // 2, N // 2)), mode="constant")
    assert np.array_equal(fr.pad_zero(PSI), q)
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    assert rel(fr.D(PSI, fk[0], np.float64, pad="zero"), PSI) < 1e-14


def test_pad_transpose_is_adjoint():
    rng = np.random.default_rng(5)
    g = rng.standard_normal((1, 2 * NZ, 2 * N))
    f = rng.standard_normal((1, NZ, N))
    assert np.isclose(np.sum(fr.pad_symmetric(f) * g), np.sum(f * fr.pad_symmetric_T(g, NZ, N)), rtol=1e-12)


def test_kernel_unit_modulus():
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    assert np.allclose(np.abs(fk * 4 * N * NZ), 1.0, atol=1e-14)


def test_z0_identity_and_constant():
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    assert rel(fr.D(PSI, fk[0], np.float64), PSI) < 1e-14
    c = np.full((NZ, N), 0.3 + 0.2j)
    assert rel(fr.D(c, fk[2], np.float64), c) < 1e-14


@pytest.mark.parametrize("px,py", [(16, 0), (0, 8), (32, 16)])
def test_cosine_grating_closed_form(px, py):
    f = ind.cosine_grating(NZ, N, px, py, 0.3)
    z = 0.1
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, [z], np.float64)
    assert rel(fr.D(f, fk[0], np.float64), ind.cosine_grating_propagated(NZ, N, px, py, 0.3, WL, VOX, z)) < 1e-13


def test_separable_equals_2d():
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    kx, ky = fr.fresnel_kernels_1d(N, NZ, WL, VOX, DIST, np.float64)
    for j in range(len(DIST)):
        assert rel(fr.D_separable(PSI, kx[j], ky[j], np.float64), fr.D(PSI, fk[j], np.float64)) < 1e-13


def test_fp32_restatement_close():
    fk32 = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float32)
    fk64 = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    for j in range(3):
        assert rel(fr.D(PSI.astype(np.complex64), fk32[j], np.float32), fr.D(PSI, fk64[j], np.float64)) < 1e-5


@pytest.mark.parametrize("j", [0, 1, 2])
def test_DT_adjoint_identity(j):
    fk = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    y = np.random.default_rng(9).standard_normal((B, NZ, N)) + 1j * np.random.default_rng(8).standard_normal((B, NZ, N))
    lhs = np.vdot(y, fr.D(PSI, fk[j], np.float64))
    rhs = np.vdot(fr.DT(y, fk[j], np.float64), PSI)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_psi_and_abs2_closed_form():
    p = np.linspace(-1, 2, 11).astype(np.float32)
    psi = fr.psi_from_proj(p, 1.5, 1e-3, np.float64)
    ph = p.astype(np.float64) * 1.5
    assert np.allclose(psi, np.exp(-ph * 1e-3) * np.exp(1j * ph), rtol=1e-12)
    assert np.allclose(fr.abs2(psi), np.exp(-2 * ph * 1e-3), rtol=1e-12)
