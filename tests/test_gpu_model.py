"""Differentiable forward model on the GPU: loss and gradient vs the FP64 oracle, finite differences."""
import numpy as np
import pytest
import torch

from oracle.model import Model

pytestmark = pytest.mark.gpu

N, NZ, NT = 64, 8, 32
WL, VOX, Z, K = 4.13e-11, 1.4e-6, 0.02, 0.5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def setup():
    from model_mosaic_amd.model import ForwardModel
    theta = np.linspace(0, np.pi, NT, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(0)
    d_true = np.abs(rng.standard_normal((NZ, N, N))).astype(np.float32)
    orc = Model(theta, N, NZ, WL, VOX, Z, K)
    data = orc.intensity(d_true, 0.02 * d_true).astype(np.float32)
    delta = (d_true + 0.1 * rng.standard_normal(d_true.shape)).astype(np.float32)
    beta = (0.02 * d_true + 0.01 * rng.standard_normal(d_true.shape)).astype(np.float32)
    fm = ForwardModel(N, NZ, theta, WL, VOX, [Z], k=K)
    return fm, orc, data, delta, beta


def test_loss_and_grad_vs_oracle(setup):
    fm, orc, data, delta, beta = setup
    dd, bb, yy = (torch.from_numpy(a).cuda() for a in (delta, beta, data))
    loss = fm.loss(dd, bb, yy).item()
    ref = orc.loss(delta.astype(np.float64), beta.astype(np.float64), data.astype(np.float64))
    gd, gb = fm.grad(dd, bb, yy)
    rgd, rgb = orc.grad(delta.astype(np.float64), beta.astype(np.float64), data.astype(np.float64))
    e_l, e_d, e_b = abs(loss - ref) / ref, rel(gd.cpu().numpy(), rgd), rel(gb.cpu().numpy(), rgb)
    print(f"model: loss {e_l:.2e}  grad delta {e_d:.2e}  grad beta {e_b:.2e}")
    assert e_l < 1e-5
    assert e_d < 1e-4 and e_b < 1e-4


def test_autograd_function(setup):
    from model_mosaic_amd.model import misfit
    fm, orc, data, delta, beta = setup
    dd = torch.from_numpy(delta).cuda().requires_grad_(True)
    bb = torch.from_numpy(beta).cuda().requires_grad_(True)
    yy = torch.from_numpy(data).cuda()
    loss = misfit(dd, bb, fm, yy)
    (2.0 * loss).backward()
    gd, gb = fm.grad(dd.detach(), bb.detach(), yy)
    assert torch.allclose(dd.grad, 2.0 * gd, rtol=1e-5, atol=1e-6 * gd.abs().max().item())
    assert torch.allclose(bb.grad, 2.0 * gb, rtol=1e-5, atol=1e-6 * gb.abs().max().item())


def test_central_finite_differences_64_voxels(setup):
    """GPU gradient vs GPU central differences on 64 random voxels (rel-L2 over them <= 1e-3)."""
    fm, orc, data, delta, beta = setup
    dd, bb, yy = (torch.from_numpy(a).cuda() for a in (delta, beta, data))
    gd, gb = (g.cpu().numpy() for g in fm.grad(dd, bb, yy))
    rng = np.random.default_rng(7)
    idx = rng.choice(delta.size, 64, replace=False)
    for which, g in ((0, gd), (1, gb)):
        fd = []
        for i in idx:
            h = 1e-2 if which == 0 else 1e-3
            vols = [dd.clone(), bb.clone()]
            vols[which].view(-1)[i] += h
            lp = fm.loss(vols[0], vols[1], yy).item()
            vols[which].view(-1)[i] -= 2 * h
            lm = fm.loss(vols[0], vols[1], yy).item()
            fd.append((lp - lm) / (2 * h))
        fd = np.array(fd)
        err = np.linalg.norm(fd - g.flat[idx]) / np.linalg.norm(g.flat[idx])
        print(f"FD check ({'delta' if which == 0 else 'beta'}): rel-L2 {err:.2e}")
        assert err < 1e-3
