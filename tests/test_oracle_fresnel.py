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


def _dct_axis_matrices(L):
    """Unnormalised DCT-II (2 cos) and DCT-III (weights 1, 2, 2, ...) matrices of length L."""
    k = np.arange(L)
    c2 = 2 * np.cos(np.pi * np.outer(k, 2 * k + 1) / (2 * L))
    c3 = np.where(k == 0, 1, 2)[None, :] * np.cos(np.pi * np.outer(2 * k + 1, k) / (2 * L))
    return c2, c3


def _makhoul_line(x, k2n, conj=False):
    """DCT-III(K * DCT-II(x)) of one line through one length-L FFT and one inverse FFT with the even / odd
    reordering v_j = x_2j, v_(L-1-j) = x_(2j+1) and, between the transforms, U_k = A_k V_k + B_k V_(L-k) with
    A_k = K_k + K_(L-k), B_k = exp(i pi k / L) (K_k - K_(L-k)) (conj(K) for DT: conj(A) and
    exp(i pi k / L) conj(K_k - K_(L-k))): the arithmetic of the build's mirror-padded Fresnel passes, restated
    independently of the oracle.  k2n is the 2L-point chirp carrying 1 / (2L)."""
    L = x.shape[-1]
    k = np.arange(L)
    v = np.concatenate([x[0::2], x[1::2][::-1]])
    V = np.fft.fft(v)
    kk, kr = k2n[:L], k2n[L - k]               # k = 0: K_L
    a, d = kk + kr, kk - kr
    c2 = np.exp(1j * np.pi * k / L)            # conj(c_k)^2, c_k = exp(-i pi k / 2L)
    if conj:
        a, b = np.conj(a), c2 * np.conj(d)
    else:
        b = c2 * d
    U = a * V + b * V[(-k) % L]
    u = np.fft.ifft(U) * L
    out = np.empty_like(u)
    out[0::2], out[1::2] = u[:L // 2], u[L // 2:][::-1]
    return out


def test_mirror_padded_propagator_is_dct_form():
    """Mirror padding makes each axis of D the product DCT-III * K * DCT-II on frame-length transforms
    (the 2N kernel is even in its index, the pad is the half-sample-symmetric extension shifted by N/2).
    Checked against the padded 2-D oracle in FP64, with explicit cosine matrices and with the one-FFT
    reordered form; a sign or conjugation slip in the reordered form fails."""
    kx, ky = fr.fresnel_kernels_1d(N, NZ, WL, VOX, DIST, np.float64)
    fker = fr.fresnel_kernel_2d(N, NZ, WL, VOX, DIST, np.float64)
    c2x, c3x = _dct_axis_matrices(N)
    c2y, c3y = _dct_axis_matrices(NZ)
    for j in range(len(DIST)):
        ref = fr.D(PSI, fker[j], np.float64)
        kxn, kyn = kx[j, :N] / (2 * N), ky[j, :NZ] / (2 * NZ)
        mats = np.einsum("yk,k,kl->yl", c3y, kyn, c2y)
        matx = np.einsum("xk,k,kl->xl", c3x, kxn, c2x)
        explicit = np.einsum("yl,blm,xm->byx", mats, PSI, matx)
        assert rel(explicit, ref) < 1e-13
        kx2, ky2 = kx[j] / (2 * N), ky[j] / (2 * NZ)
        rows = np.apply_along_axis(_makhoul_line, 2, PSI, kx2)
        both = np.apply_along_axis(_makhoul_line, 1, rows, ky2)
        assert rel(both, ref) < 1e-13
        # DT of the mirror-padded propagator is the same pipeline with conj(K)
        conj_rows = np.apply_along_axis(_makhoul_line, 2, PSI, kx2, True)
        conj_both = np.apply_along_axis(_makhoul_line, 1, conj_rows, ky2, True)
        assert rel(conj_both, fr.DT(PSI, fker[j], np.float64)) < 1e-13
