"""HIP Fresnel D / DT parity against the oracle and its golden fixtures (needs a GPU)."""
import os

import numpy as np
import pytest
import torch

from oracle import fresnel as fr
from oracle import independent as ind

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLD, "fresnel_small.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def prop(g):
    from model_mosaic_amd.processing.propagation import Propagation
    nt, nz, n = g["psi"].shape
    return Propagation(n, nz, nt, len(g["dist"]), float(g["wavelength"]), float(g["voxelsize"]), g["dist"])


def test_D_DT_golden(g, prop):
    x = torch.from_numpy(g["psi"]).cuda()
    for j in range(len(g["dist"])):
        assert rel(prop.D(x, j).cpu().numpy(), g["D"][j]) < 1e-5
        assert rel(prop.DT(x, j).cpu().numpy(), g["DT"][j]) < 1e-5


def test_D_vs_fp32_reference_restatement(g, prop):
    n, nz = g["psi"].shape[2], g["psi"].shape[1]
    fk = fr.fresnel_kernel_2d(n, nz, float(g["wavelength"]), float(g["voxelsize"]), g["dist"], np.float32)
    x = torch.from_numpy(g["psi"]).cuda()
    for j in range(3):
        assert rel(prop.D(x, j).cpu().numpy(), fr.D(g["psi"], fk[j], np.float32)) < 1e-5


def test_z0_identity_2d_input_and_inplace(g, prop):
    x = torch.from_numpy(g["psi"]).cuda()
    assert rel(prop.D(x[1], 0).cpu().numpy(), g["psi"][1]) < 1e-6
    y = x.clone()
    prop.D(y, 2, out=y)
    assert rel(y.cpu().numpy(), g["D"][2]) < 1e-5


@pytest.mark.parametrize("j", [1, 2])
def test_adjoint_identity(g, prop, j):
    """<b, D a> = <DT b, a>, inner products in FP64.

    For random b the sum <b, D a> cancels down to ~1/sqrt(N) of ||b|| ||D a||,
    so FP32 rounding of D and DT (~2e-7 per element against the FP64 operator)
    shows up magnified in |lhs - rhs| / |lhs|; that ratio gets 5e-6.  The error
    is also bounded at the Cauchy-Schwarz scale ||b|| ||D a|| and on the
    non-cancelling pair b = D a, both at 1e-6 (a wrong adjoint gives O(1))."""
    rng = np.random.default_rng(3)
    a = torch.from_numpy(g["psi"]).cuda()
    b = torch.from_numpy((rng.standard_normal(g["psi"].shape) + 1j * rng.standard_normal(g["psi"].shape))
                         .astype(np.complex64)).cuda()
    c128 = lambda t: t.flatten().to(torch.complex128)
    da = prop.D(a, j)
    lhs = torch.vdot(c128(b), c128(da)).item()
    rhs = torch.vdot(c128(prop.DT(b, j)), c128(a)).item()
    cs = abs(lhs - rhs) / (c128(b).norm().item() * c128(da).norm().item())
    lhs2 = torch.vdot(c128(da), c128(da)).item()
    rhs2 = torch.vdot(c128(prop.DT(da, j)), c128(a)).item()
    print(f"adjoint j={j}: |lhs-rhs|/|lhs| {abs(lhs - rhs) / abs(lhs):.2e}  cauchy-schwarz {cs:.2e}  "
          f"b=Da {abs(lhs2 - rhs2) / abs(lhs2):.2e}")
    assert abs(lhs - rhs) / abs(lhs) < 5e-6
    assert cs < 1e-6
    assert abs(lhs2 - rhs2) / abs(lhs2) < 1e-6


def test_cosine_grating_closed_form():
    from model_mosaic_amd.processing.propagation import Propagation
    n, nz, wl, vox, z = 256, 128, 4.1333e-11, 0.69e-6, 0.8
    p = Propagation(n, nz, 1, 1, wl, vox, [z])
    f = ind.cosine_grating(nz, n, 32, 16, 0.3).astype(np.complex64)
    got = p.D(torch.from_numpy(f).cuda(), 0).cpu().numpy()
    assert rel(got, ind.cosine_grating_propagated(nz, n, 32, 16, 0.3, wl, vox, z)) < 1e-5


@pytest.mark.parametrize("n,nz", [(128, 64), (512, 1024), (2048, 256),
                                  # 3 * 2^k frames (radix-12 FFT passes): mosaic tiles of 96 rows, widths 1536 / 3072
                                  (96, 48), (1536, 96), (3072, 96), (64, 24), (384, 768)])
def test_sizes_vs_fp64_separable(n, nz):
    from model_mosaic_amd.processing.propagation import Propagation
    wl, vox, dist = 4.1333e-11, 0.69e-6, [0.05, 0.8]
    rng = np.random.default_rng(n)
    psi = np.exp(1j * rng.uniform(0, 2, (2, nz, n))).astype(np.complex64)
    p = Propagation(n, nz, 2, 2, wl, vox, dist)
    kx, ky = fr.fresnel_kernels_1d(n, nz, wl, vox, dist, np.float64)
    for j in range(2):
        assert rel(p.D(torch.from_numpy(psi).cuda(), j).cpu().numpy(), fr.D_separable(psi, kx[j], ky[j], np.float64)) < 1e-5
    # DT is the adjoint of D at every size (FP64 inner products)
    a = torch.from_numpy(psi).cuda()
    b = torch.from_numpy(np.exp(1j * rng.uniform(0, 2, (2, nz, n))).astype(np.complex64)).cuda()
    lhs = torch.vdot(p.D(a, 1).flatten().to(torch.complex128), b.flatten().to(torch.complex128)).item()
    rhs = torch.vdot(a.flatten().to(torch.complex128), p.DT(b, 1).flatten().to(torch.complex128)).item()
    assert abs(lhs - rhs) / (a.norm().item() * b.norm().item()) < 1e-6


def test_zero_padding_mode():
    from model_mosaic_amd.processing.propagation import Propagation
    n, nz, wl, vox = 64, 32, 4.13e-11, 1.4e-6
    psi = (np.random.default_rng(1).standard_normal((2, nz, n)) + 0j).astype(np.complex64)
    p = Propagation(n, nz, 2, 1, wl, vox, [0.1], pad="zero")
    fk = fr.fresnel_kernel_2d(n, nz, wl, vox, [0.1], np.float64)
    assert rel(p.D(torch.from_numpy(psi).cuda(), 0).cpu().numpy(), fr.D(psi, fk[0], np.float64, pad="zero")) < 1e-5


def test_large_parity_and_validation(g):
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.processing.propagation_large import PropagationLarge
    nt, nz, n = g["psi"].shape
    wl, vox, dist = float(g["wavelength"]), float(g["voxelsize"]), g["dist"]
    ref_p = Propagation(n, nz, nt, len(dist), wl, vox, dist)
    lrg = PropagationLarge(n, nz, wl, vox, dist)
    for j in range(len(dist)):
        ref = ref_p.D(torch.from_numpy(g["psi"]).cuda(), j).cpu().numpy()
        for chunks in [(32, 128), (16, 32), (8, 16), (4, 8)]:
            got = lrg.D(g["psi"], j, list(chunks))
            assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-3
    small = PropagationLarge(32, 16, 4.13e-11, 1.4e-6, [0.0])
    psi = np.zeros((1, 16, 32), dtype=np.complex64)
    for chunks in [(5, 16), (8, 9)]:
        with pytest.raises(AssertionError):
            small.D(psi, 0, list(chunks))


def test_psi_and_abs2():
    from model_mosaic_amd.step_kernels import abs2, psi_from_proj
    p = np.random.default_rng(2).uniform(-2, 3, (4, 16, 32)).astype(np.float32)
    psi = psi_from_proj(torch.from_numpy(p).cuda(), np.float32(0.7), np.float32(1e-3)).cpu().numpy()
    ref = fr.psi_from_proj(p, 0.7, 1e-3, np.float64)
    assert rel(psi, ref) < 1e-6
    assert rel(abs2(torch.from_numpy(psi).cuda()).cpu().numpy(), fr.abs2(ref)) < 1e-6


def test_large_streams_frame_batches_bitwise(g):
    """A GPU budget of two frames makes PropagationLarge stream 5 frames in 3 batches through the three-stream
    pipeline: D and DT equal the one-batch Propagation bit for bit (frames are independent)."""
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.processing.propagation_large import PropagationLarge
    _, nz, n = g["psi"].shape
    wl, vox, dist = float(g["wavelength"]), float(g["voxelsize"]), g["dist"]
    rng = np.random.default_rng(4)
    psi = np.exp(1j * rng.uniform(0, 2, (5, nz, n))).astype(np.complex64)
    lrg = PropagationLarge(n, nz, wl, vox, dist, gpu_budget_bytes=2 * 4 * nz * n * 8 + 1)
    assert lrg._batch(5) == 2
    ref = Propagation(n, nz, 5, len(dist), wl, vox, dist)
    for j in range(len(dist)):
        assert np.array_equal(lrg.D(psi, j, [nz, 2 * n]), ref.D(torch.from_numpy(psi).cuda(), j).cpu().numpy())
        assert np.array_equal(lrg.DT(psi, j, [nz, 2 * n]), ref.DT(torch.from_numpy(psi).cuda(), j).cpu().numpy())


@pytest.mark.parametrize("budget_frames", [None, 2])
def test_large_banded_psi_buffer_filled_band_by_band(g, budget_frames):
    """step3_propagation_large.py:210-226: psi_buffer is a BandedPinned along nz that the caller fills band by
    band (.real / .imag per band from a projection slab) and hands to D, which returns a BandedPinned: the
    result equals Propagation.D of the whole array, in one device batch and streamed in frame ranges."""
    from model_mosaic_amd.processing.pipeline import BandedPinned
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.processing.propagation_large import PropagationLarge
    _, nz, n = g["psi"].shape
    wl, vox, dist = float(g["wavelength"]), float(g["voxelsize"]), g["dist"]
    nt = 3
    kw = {} if budget_frames is None else {"gpu_budget_bytes": budget_frames * 4 * nz * n * 8 + 1}
    lrg = PropagationLarge(n, nz, wl, vox, dist, **kw)
    proj = np.random.default_rng(8).uniform(0, 2, (nt, nz, n)).astype(np.float32)
    buf = lrg.psi_buffer(nt)
    assert isinstance(buf, BandedPinned) and buf.n_bands >= 8 and buf.shape == (nt, nz, n)
    for bi, band in enumerate(buf.bands):
        z0 = bi * buf.band_rows
        ph = proj[:, z0:z0 + buf.band_rows]
        at = np.exp(-ph * 1e-3).astype(np.float32)
        band.real = at * np.cos(ph)
        band.imag = at * np.sin(ph)
    psi = np.asarray(buf)
    out = lrg.D(buf, 1, [nz, 2 * n])
    assert isinstance(out, BandedPinned) and out.shape == (nt, nz, n)
    ref = Propagation(n, nz, nt, len(dist), wl, vox, dist).D(torch.from_numpy(psi).cuda(), 1).cpu().numpy()
    assert np.array_equal(np.asarray(out), ref)
    assert np.array_equal(out[:, :buf.band_rows], ref[:, :buf.band_rows])   # one band: a view of its block
