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
    the taps of its item's entries that land on it (lists decoded as the kernel does, against a direct enumeration),
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
def test_fbp_filter_weights_equal_the_oracle(n):
    """Host weights of the product (FBPFilter.calc_filter) = the oracle's restatement, bit for bit."""
    from model_mosaic_amd.processing.fbp_filter import FILTERS, calc_filter_host

    from oracle import fbp
    for name in FILTERS:
        assert np.array_equal(calc_filter_host(n, name), fbp.calc_filter(n, name)), name


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
