"""HIP Paganin phase retrieval parity against the oracle (needs a GPU)."""
import numpy as np
import pytest
import torch

from oracle import paganin as pg

pytestmark = pytest.mark.gpu

WL, VOX, ALPHA = 4.13e-11, 1.4e-6, 1e-3


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n,nz,b,ntheta,dist", [(64, 32, 3, 3, 0.1), (256, 128, 4, 6, 1.0), (96, 48, 5, 5, 0.02),
                                                (2048, 64, 2, 2, 0.5), (1536, 96, 1, 2, 0.3), (16, 16, 7, 8, 1.0)])
def test_retrieve_vs_fp64(n, nz, b, ntheta, dist):
    """Odd and even frame counts (two frames per complex transform), batches below ntheta, 3 * 2^k sizes."""
    from model_mosaic_amd.processing.paganin import Paganin
    x = np.random.default_rng(n + b).uniform(0.5, 1.0, (b, nz, n)).astype(np.float32)
    p = Paganin(n, nz, ntheta, WL, VOX, dist, ALPHA)
    got = p.retrieve(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = pg.retrieve(x, WL, VOX, dist, ALPHA)
    err = rel(got, ref)
    print(f"paganin n={n} nz={nz} b={b}: rel-L2 vs FP64 {err:.2e}")
    assert err < 1e-5
    assert rel(got, pg.retrieve(x, WL, VOX, dist, ALPHA, np.float32)) < 1e-5


def test_single_mode_closed_form_and_2d_input():
    from model_mosaic_amd.processing.paganin import Paganin
    n, nz, p_, q_, dist = 128, 64, 3, 7, 1.0
    y, x = np.mgrid[0:nz, 0:n]
    phase = 2 * np.pi * (p_ * y / nz + q_ * x / n)
    inten = (1.0 + 0.3 * np.cos(phase)).astype(np.float32)
    h = pg.mode_gain(p_, q_, n, nz, WL, VOX, dist, ALPHA)
    got = Paganin(n, nz, 1, WL, VOX, dist, ALPHA).retrieve(torch.from_numpy(inten).cuda()).cpu().numpy()
    assert got.shape == (nz, n)
    assert np.max(np.abs(got - (-np.log(1.0 + 0.3 * h * np.cos(phase))))) < 2e-6


def test_clip_and_errors():
    from model_mosaic_amd.processing.paganin import Paganin
    p = Paganin(32, 16, 2, WL, VOX, 0.1, ALPHA)
    z = p.retrieve(torch.zeros((2, 16, 32), device="cuda")).cpu().numpy()
    assert np.allclose(z, -np.log(1e-6), rtol=1e-6)
    with pytest.raises(AssertionError):
        p.retrieve(torch.ones((3, 16, 32), device="cuda"))
    with pytest.raises(ValueError):
        p.retrieve(torch.ones((2, 16, 16), device="cuda"))


def test_paganin_vs_large():
    """tests/test_large_variants.py:227-257: PaganinLarge over several chunkings = Paganin (rel-max 1e-4)."""
    from model_mosaic_amd.processing.paganin import Paganin
    from model_mosaic_amd.processing.paganin_large import PaganinLarge
    n, nz, ntheta = 64, 32, 3
    x = (0.5 + 0.5 * np.random.default_rng(0).random((ntheta, nz, n))).astype(np.float32)
    for dist in [0.02, 0.1, 1.0]:
        ref = Paganin(n, nz, ntheta, WL, VOX, dist, ALPHA).retrieve(torch.from_numpy(x).cuda()).cpu().numpy()
        lrg = PaganinLarge(n, nz, WL, VOX, dist, ALPHA)
        for chunks in [(nz, n), (16, 32), (8, 16), (4, 8)]:
            got = np.asarray(lrg.retrieve(x, list(chunks)))
            assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-4
        assert rel(lrg.retrieve(x[1], [nz, n]), ref[1]) < 1e-6
    with pytest.raises(AssertionError):
        lrg.retrieve(x, [5, n])
    with pytest.raises(AssertionError):
        lrg.retrieve(x, [nz, 7])


def test_large_streams_frame_batches():
    """A GPU budget of two frames makes PaganinLarge stream 5 frames in 3 batches through the three-stream
    pipeline; frames are paired differently than in one batch, so the match is to FP32 rounding."""
    from model_mosaic_amd.processing.paganin import Paganin
    from model_mosaic_amd.processing.paganin_large import PaganinLarge
    n, nz = 64, 32
    x = (0.5 + 0.5 * np.random.default_rng(2).random((5, nz, n))).astype(np.float32)
    lrg = PaganinLarge(n, nz, WL, VOX, 0.1, ALPHA, gpu_budget_bytes=2 * 5 * nz * n * 4 + 1)
    assert lrg._batch(5) == 2
    ref = Paganin(n, nz, 5, WL, VOX, 0.1, ALPHA).retrieve(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel(lrg.retrieve(x, [nz, n]), ref) < 1e-6
