"""The dxchange HDF5 layer feeding the device operators (needs a GPU): slabs streamed from a file through
page-locked buffers give the same results as the arrays they were written from; a step2-style loop
(volume slabs -> R -> sinogram file banked along z) round-trips."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dx():
    from model_mosaic_amd.iohdf5 import dxchange_hdf5_chunks as mod
    if not mod.LIB_PATH.exists():
        pytest.skip("libmosaic_h5.so not built (no HDF5 C library)")
    return mod


def test_projection_slabs_to_device_and_fresnel(dx, tmp_path):
    from model_mosaic_amd.processing.propagation import Propagation
    nt, nz, n = 7, 32, 64
    f = str(tmp_path / "proj.h5")
    ctx = dx.tomo_initx(f, (nt, nz, n), np.float32, vchunks=(4, nz, n), stype="proj", nbanks=2)
    p = np.random.default_rng(3).uniform(0, 2, (nt, nz, n)).astype(np.float32)
    for iv in range(2):
        dx.tomo_writex(f, p[4 * iv:4 * iv + 4], ivchunk=(iv, 0, 0), ctx=ctx)
    prop = Propagation(n, nz, nt, 1, 4.13e-11, 1.4e-6, [0.05])
    whole = prop.D(torch.polar(torch.ones(nt, nz, n, device="cuda"), torch.from_numpy(p).cuda()), 0)
    seen = 0
    for start, slab in dx.slabs_to_device(f, 0, 3, torch.device("cuda", 0)):
        assert torch.equal(slab, torch.from_numpy(p[start:start + slab.shape[0]]).cuda())
        d = prop.D(torch.polar(torch.ones_like(slab), slab), 0)
        assert torch.equal(d, whole[start:start + slab.shape[0]])
        seen += slab.shape[0]
    assert seen == nt


def test_step2_style_volume_to_sinogram_file(dx, tmp_path):
    """Volume slabs along z read from a file, R on the device, the sinogram slab written into a file banked
    along z (as step2_radon.py:147-200 writes proj.h5), read back: equals R of the whole volume."""
    from model_mosaic_amd.processing.tomo import TomoReal
    nz, n, nt = 8, 64, 24
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    vol = np.random.default_rng(4).standard_normal((nz, n, n)).astype(np.float32)
    fin, fout = str(tmp_path / "obj.h5"), str(tmp_path / "proj.h5")
    cin = dx.tomo_initx(fin, (nz, n, n), np.float32, vchunks=(4, n, n), stype="proj", nbanks=1)
    for iv in range(2):
        dx.tomo_writex(fin, vol[4 * iv:4 * iv + 4], ivchunk=(iv, 0, 0), ctx=cin)
    cout = dx.tomo_initx(fout, (nt, nz, n), np.float32, vchunks=(nt, 4, n), stype="slice", nbanks=2)
    dx.write_theta(fout, np.rad2deg(th))
    tomo = TomoReal(n, 4, th)
    for z0, slab in dx.slabs_to_device(fin, 0, 4, torch.device("cuda", 0)):
        sino = tomo.R(slab).cpu().numpy()                       # (nt, 4, n)
        dx.tomo_writex(fout, sino, ivchunk=(0, z0 // 4, 0), ctx=cout)
    got = dx.read_box(fout, (0, 0, 0), (nt, nz, n))
    ref = TomoReal(n, nz, th).R(torch.from_numpy(vol).cuda()).cpu().numpy()
    assert np.array_equal(got, ref)
    assert np.allclose(np.deg2rad(dx.read_theta(fout)), th, rtol=0, atol=1e-6)
