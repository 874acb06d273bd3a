"""Host-side logic that needs no GPU: chunk validation, status mapping, phantoms."""
import numpy as np
import pytest

from model_mosaic_amd import _lib, phantom
from model_mosaic_amd.processing.propagation_large import check_prop_chunks
from model_mosaic_amd.processing.tomo_large import check_tomo_chunks


def test_prop_chunk_validation_matches_reference():
    # tests/test_large_variants.py:205-223: non-dividing chunks raise AssertionError
    check_prop_chunks(16, 32, (1, 16, 32), (8, 16))
    with pytest.raises(AssertionError, match="CHUNK_NZ=5"):
        check_prop_chunks(16, 32, (1, 16, 32), (5, 16))
    with pytest.raises(AssertionError, match="CHUNK_2N=9"):
        check_prop_chunks(16, 32, (1, 16, 32), (8, 9))
    with pytest.raises(AssertionError, match="mismatches propagator"):
        check_prop_chunks(16, 32, (1, 16, 30), (8, 16))


def test_tomo_chunk_validation():
    for chunks in [(64, 24, 64), (32, 12, 32), (16, 6, 16), (8, 6, 8), (64, 24, 128)]:
        check_tomo_chunks(64, 24, chunks)
    with pytest.raises(AssertionError, match="CHUNK_N=24"):
        check_tomo_chunks(64, 24, (24, 12, 32))
    with pytest.raises(AssertionError, match="CHUNK_THETA=5"):
        check_tomo_chunks(64, 24, (32, 5, 32))
    with pytest.raises(AssertionError):
        check_tomo_chunks(64, 24, (32, 12))


def test_status_mapping():
    assert _lib.check(_lib.MM_OK, "x") is None
    with pytest.raises(ValueError):
        _lib.check(_lib.MM_ERR_ARG, "x")
    with pytest.raises(MemoryError):
        _lib.check(_lib.MM_ERR_OOM, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.MM_ERR_UNSUPPORTED, "x")
    with pytest.raises(_lib.MosaicLibraryError):
        _lib.check(_lib.MM_ERR_HIP, "x")


def test_phantom_seeded_and_bounded():
    a = phantom.ellipsoids(8, 64, seed=3)
    b = phantom.ellipsoids(8, 64, seed=3)
    assert a.shape == (8, 64, 64) and bool((a == b).all())
    assert float(a.max()) <= 12 * 1e-6 and float(a.min()) >= 0 and float(a.max()) > 0
    th = phantom.angles(384)
    assert th.dtype == np.float32 and th[0] == 0 and th[-1] < np.pi


@pytest.mark.parametrize("n", [16, 32, 64, 96, 128, 192, 256, 384, 768, 1536])
def test_radon_plan_tables_match_kernel_indexing(n):
    """Host-only: every gather staging offset lies inside a half-spectrum plane (tiles at the right
    edge of small grids reach past 2n and must wrap), every patch element is staged once, every tap
    read stays inside the LDS patch, every scatter slot names a staged entry, every written scatter point gets exactly
    the taps of its item's entries that land on it (lists decoded as the kernel does, against a direct enumeration;
    likewise the matrix-core scatter's anchor-class runs and window test, with its row map and sample-order staging),
    the column passes cover every row the gather reads and read no row the scatter leaves unwritten (dense and sparse
    angle sets)."""
    import ctypes as C

    import numpy as np

    from model_mosaic_amd import _lib
    lib = _lib.load()
    for nt, span in ((7, np.pi), (24, 2 * np.pi), (37, np.pi), (400, np.pi)):
        th = np.linspace(0, span, nt, endpoint=False).astype(np.float32).astype(np.float64)
        rc = lib.mm_radon_plan_check(n, nt, th.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == _lib.MM_OK, _lib.last_error()


@pytest.mark.parametrize("n", [16, 64, 96, 2048])
def test_fbp_filter_weights_match_the_oracle(n):
    """Host weights of the product (FBPFilter.calc_filter, its own quadrature construction) agree with the
    oracle's restatement of the reference's weights.  The reference builds its quadrature through an inverse
    Vandermonde matrix on nodes starting at 1e-40, which carries ~1e-9 of rounding (the product's rule is exact
    to 1e-15, see the moment test); after the float32 cast the two agree to a few float32 ulps."""
    from model_mosaic_amd.processing.fbp_filter import FILTERS, calc_filter_host

    from oracle import fbp
    for name in FILTERS:
        a, b = calc_filter_host(n, name).astype(np.float64), fbp.calc_filter(n, name).astype(np.float64)
        assert np.max(np.abs(a - b)) <= 2e-6 * np.max(np.abs(b)), name


@pytest.mark.parametrize("N", [25, 33, 40])
def test_ramp_quadrature_moments_exact(N):
    """The product's ramp rule integrates t * t^d over [0, T] exactly for d <= 11 (no taper for N <= 40):
    a wrong window, gap integral or overlap share is an O(1) error."""
    from model_mosaic_amd.processing.fbp_filter import ramp_quadrature
    t = np.arange(N) / (4.0 * (N - 1))
    w = ramp_quadrature(t)
    T = t[-1]
    for d in range(12):
        exact = T ** (d + 2) / (d + 2)
        assert abs(np.sum(w * t ** d) - exact) < 1e-13 * exact, d


def test_fbp_unknown_filter_name():
    from model_mosaic_amd.processing.fbp_filter import calc_filter_host
    with pytest.raises(ValueError):
        calc_filter_host(16, "gaussian")


def test_out_buffer_validation():
    """check_out rejects an output buffer of the wrong size, dtype, device or layout before any address
    reaches a kernel."""
    import torch

    from model_mosaic_amd._carriers import check_out
    cpu = torch.device("cpu")
    good = torch.empty((4, 3, 8), dtype=torch.float32)
    assert check_out(good, (4, 3, 8), torch.float32, cpu, "t") is good
    assert check_out(good.view(12, 8), (4, 3, 8), torch.float32, cpu, "t") is not None
    with pytest.raises(ValueError, match="shape"):
        check_out(torch.empty((4, 3, 7)), (4, 3, 8), torch.float32, cpu, "t")
    with pytest.raises(ValueError, match="dtype"):
        check_out(torch.empty((4, 3, 8), dtype=torch.complex64), (4, 3, 8), torch.float32, cpu, "t")
    with pytest.raises(ValueError, match="contiguous"):
        check_out(torch.empty((8, 3, 4)).transpose(0, 2), (4, 3, 8), torch.float32, cpu, "t")
    with pytest.raises(ValueError, match="expected cuda:0"):
        check_out(good, (4, 3, 8), torch.float32, torch.device("cuda", 0), "t")
    with pytest.raises(ValueError, match="torch tensor"):
        check_out(np.empty((4, 3, 8), np.float32), (4, 3, 8), torch.float32, cpu, "t")


@pytest.mark.parametrize("n,name", [(16, "ramp"), (64, "shepp"), (96, "hann"), (512, "parzen"), (2048, "shepp")])
def test_fbp_two_n_tables_reproduce_the_padded_filter(n, name):
    """Host-only: the length-2n form (two_n_tables: G, A, B) applied in FP64 with numpy FFTs equals the
    oracle's edge-padded length-4n filter (the reference's algorithm) to rounding."""
    from model_mosaic_amd.processing.fbp_filter import calc_filter_host, two_n_tables

    from oracle import fbp
    w = calc_filter_host(n, name)
    G, A, B = (a.astype(np.float64) for a in two_n_tables(w, n))
    x = np.random.default_rng(n).standard_normal((5, n))
    L = 2 * n
    core = np.fft.irfft(np.fft.rfft(x, L) * (G * L), L)[:, :n]
    y = core + x[:, :1] * A + x[:, -1:] * B
    ref = fbp.filter(x, w.astype(np.float64))
    # G, A, B are handed over in float32: 1e-6 relative
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-6


def test_banded_pinned_semantics():
    """BandedPinned (processing/pipeline.py:92-252 interface) in pageable memory: band layout, reads inside one
    band are views, reads across bands are copies, writes of scalars and arrays, band-wise copy_to / copy_from,
    the band-count rule, and the errors the reference raises."""
    from model_mosaic_amd.processing.pipeline import BandedPinned, pick_n_bands
    ref = (np.arange(2 * 12 * 5) + 1j * np.arange(2 * 12 * 5)[::-1]).reshape(2, 12, 5).astype(np.complex64)
    bp = BandedPinned(ref.shape, np.complex64, n_bands=4, band_axis=1, pinned=False)
    assert bp.band_rows == 3 and len(bp.bands) == 4 and bp.bands[0].shape == (2, 3, 5) and bp.nbytes == ref.nbytes
    for i, band in enumerate(bp.bands):          # the band-wise fill of step3_propagation_large.py:210-226
        band.real = ref[:, 3 * i:3 * i + 3].real
        band.imag = ref[:, 3 * i:3 * i + 3].imag
    assert np.array_equal(np.asarray(bp), ref)
    v = bp[:, 3:6]
    assert np.array_equal(v, ref[:, 3:6]) and np.shares_memory(v, bp.bands[1])
    assert np.array_equal(bp[1, 2:10, 1:4], ref[1, 2:10, 1:4])
    assert np.array_equal(bp[:, 7], ref[:, 7]) and np.array_equal(bp[0, -1], ref[0, -1])
    assert np.array_equal(bp[1], ref[1])
    bp[:, 4:8] = 0
    want = ref.copy()
    want[:, 4:8] = 0
    assert np.array_equal(np.asarray(bp), want)
    bp[0, 2:11, :2] = ref[0, 2:11, :2] * 2
    want[0, 2:11, :2] = ref[0, 2:11, :2] * 2
    assert np.array_equal(np.asarray(bp), want)
    dst = np.zeros((1, 12, 5), np.complex64)
    bp.copy_to(dst, (slice(1, 2),))
    assert np.array_equal(dst[0], want[1])
    bp.copy_from(ref[:1], (slice(0, 1),))
    want[0] = ref[0]
    assert np.array_equal(np.asarray(bp), want)
    assert np.array_equal(np.abs(bp - want), np.zeros_like(want, dtype=np.float32))   # ndarray ops via __array__
    bp.fill(3)
    assert np.all(np.asarray(bp) == 3)
    with pytest.raises(ValueError):
        BandedPinned((2, 12, 5), np.float32, n_bands=5, pinned=False)
    with pytest.raises(NotImplementedError):
        bp[:, ::2]
    assert pick_n_bands((4, 96, 64), np.complex64, 1, 8) == 8
    assert pick_n_bands((4, 96, 64), np.complex64, 1, 8, max_band_bytes=4 * 6 * 64 * 8) == 16
    assert pick_n_bands((1, 7, 3), np.float32, 1, 2) == 7
