"""Uneven mosaic batch (R, psi, Fresnel D, |.|^2 per tile) as one HIP graph (needs a GPU)."""
import numpy as np
import pytest
import torch

from oracle import fresnel as fr
from oracle import usfft

pytestmark = pytest.mark.gpu

WL, VOX, Z = 1.24e-9 / 30.0, 1.38e-6, 0.2


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def tiles_and_inputs():
    """Six tiles of widths 96 / 128 / 192 / 256 (two pairs share geometry), 8-24 rows, 33-66 angles."""
    th33 = np.linspace(0, np.pi, 33, endpoint=False).astype(np.float32)
    th66 = np.linspace(0, np.pi, 66, endpoint=False).astype(np.float32)
    th50 = np.sort(np.random.default_rng(9).uniform(0, np.pi, 50)).astype(np.float32)
    tiles = [(96, 24, th33), (128, 16, th66), (192, 8, th33), (96, 24, th33), (128, 8, th66), (256, 16, th50)]
    from model_mosaic_amd import phantom
    deltas = [phantom.ellipsoids(nz, n, seed=20 + t, device="cuda") for t, (n, nz, _) in enumerate(tiles)]
    return tiles, deltas


def k_for(deltas, tiles):
    """k such that the largest phase k p is about 2 rad (p = R delta, |p| <= ~ n max(delta) / sqrt(n ntheta))."""
    n, _, th = tiles[0]
    return 2.0 / float(deltas[0].max()) / np.sqrt(n / len(th)) / 0.5


def eager_chain(tiles, deltas, k, beta):
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.processing.tomo import TomoReal
    from model_mosaic_amd.step_kernels import abs2, psi_from_proj
    outs = []
    for (n, nz, th), d in zip(tiles, deltas):
        p = TomoReal(n, nz, th).R(d)
        psi = psi_from_proj(p, np.float32(k), np.float32(beta))
        u = Propagation(n, nz, len(th), 1, WL, VOX, [Z]).D(psi, 0)
        outs.append(abs2(u))
    return outs


def test_graph_equals_per_tile_chain_and_replays():
    from model_mosaic_amd.mosaic import MosaicBatch
    tiles, deltas = tiles_and_inputs()
    k, beta = k_for(deltas, tiles), 1e-3
    ref = eager_chain(tiles, deltas, k, beta)
    mb = MosaicBatch(tiles, WL, VOX, Z, k, beta)
    info = mb.info()
    assert info["radon_plans"] == 4 and info["fresnel_plans"] == 5 and info["tiles"] == 6
    outs = mb.out_buffers()
    for rep in range(3):
        for o in outs:
            o.fill_(-1.0)
        mb.run(deltas, outs)
        torch.cuda.synchronize()
        for t, (o, r) in enumerate(zip(outs, ref)):
            assert torch.equal(o, r), f"replay {rep} tile {t}"
        _ = torch.randn(1 << 20, device="cuda").sum().item()   # eager work between replays
    assert mb.info()["graph_captures"] == 1
    outs2 = mb.run(deltas)            # new output pointers: captured again
    torch.cuda.synchronize()
    assert mb.info()["graph_captures"] == 2
    for o, r in zip(outs2, ref):
        assert torch.equal(o, r)


def test_no_graph_and_autotune_same_results():
    from model_mosaic_amd.mosaic import MosaicBatch
    tiles, deltas = tiles_and_inputs()
    k, beta = k_for(deltas, tiles), 1e-3
    a = MosaicBatch(tiles, WL, VOX, Z, k, beta, graph=False).run(deltas)
    mb = MosaicBatch(tiles, WL, VOX, Z, k, beta, autotune=True)
    b = mb.run(deltas)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    info = mb.info()
    assert all(g in (1, 2, 4, 8, 16) for g in info["gather_xcd_group"])
    assert all(r in (32, 64) for r in info["gather_root"]) and all(z in (1, 2) for z in info["gather_zp"])


def test_branches_same_results():
    """The tiles split over 1, 2 or 3 concurrent graph branches (each with its own workspaces): same bits, and
    every branch gets tiles."""
    from model_mosaic_amd.mosaic import MosaicBatch
    tiles, deltas = tiles_and_inputs()
    k, beta = k_for(deltas, tiles), 1e-3
    ref = MosaicBatch(tiles, WL, VOX, Z, k, beta, branches=1).run(deltas)
    for nb in (2, 3):
        mb = MosaicBatch(tiles, WL, VOX, Z, k, beta, branches=nb)
        for rep in range(2):
            out = mb.run(deltas)
            torch.cuda.synchronize()
            for t, (o, r) in enumerate(zip(out, ref)):
                assert torch.equal(o, r), f"branches {nb} replay {rep} tile {t}"
        info = mb.info()
        assert info["branches"] == nb and sorted(set(info["tile_branch"])) == list(range(nb))
        mb.free()
    with pytest.raises(ValueError):
        MosaicBatch(tiles, WL, VOX, Z, k, beta, branches=9)


def test_tile_vs_fp64_oracle():
    """Tile 2 (n=192, 8 rows, 33 angles): intensity against R, psi, D, |.|^2 of the FP64 oracle."""
    from model_mosaic_amd.mosaic import MosaicBatch
    tiles, deltas = tiles_and_inputs()
    k, beta = k_for(deltas, tiles), 1e-3
    out = MosaicBatch(tiles, WL, VOX, Z, k, beta).run(deltas)[2].cpu().numpy()
    n, nz, th = tiles[2]
    d = deltas[2].cpu().numpy()
    p = usfft.tomo_R(d, th, np.float64)
    psi = fr.psi_from_proj(p, k, beta, np.float64)
    fk = fr.fresnel_kernel_2d(n, nz, WL, VOX, [Z], np.float64)
    ref = fr.abs2(fr.D(psi, fk[0], np.float64))
    err = rel(out, ref)
    print(f"mosaic tile n={n}: intensity rel-L2 vs FP64 oracle {err:.2e}")
    assert err < 1e-5


def test_bad_inputs_rejected():
    from model_mosaic_amd.mosaic import MosaicBatch
    tiles, deltas = tiles_and_inputs()
    mb = MosaicBatch(tiles, WL, VOX, Z, 1.0, 1e-3)
    with pytest.raises(ValueError):
        mb.run(deltas[:-1])
    bad = list(deltas)
    bad[1] = bad[1][:, :, :64].contiguous()
    with pytest.raises(ValueError):
        mb.run(bad)
    with pytest.raises(ValueError):
        MosaicBatch([(96, 12, np.zeros(4, np.float32))], WL, VOX, Z, 1.0, 1e-3)   # 12 rows: Fresnel needs >= 24 for 3 * 2^k
