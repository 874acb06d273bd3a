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


@pytest.fixture(scope="module", params=["symmetric", "zero"])
def setup(request):
    """Mirror padding runs the element-wise steps inside the Fresnel passes (DCT form); zero padding runs
    them as kernels of their own around D and DT."""
    from model_mosaic_amd.model import ForwardModel
    pad = request.param
    theta = np.linspace(0, np.pi, NT, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(0)
    d_true = np.abs(rng.standard_normal((NZ, N, N))).astype(np.float32)
    orc = Model(theta, N, NZ, WL, VOX, Z, K, pad=pad)
    data = orc.intensity(d_true, 0.02 * d_true).astype(np.float32)
    delta = (d_true + 0.1 * rng.standard_normal(d_true.shape)).astype(np.float32)
    beta = (0.02 * d_true + 0.01 * rng.standard_normal(d_true.shape)).astype(np.float32)
    fm = ForwardModel(N, NZ, theta, WL, VOX, [Z], k=K, pad=pad)
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
    # the gradients are sums with heavy cancellation (random residuals); FP32 rounding against the FP64 oracle
    # measures 2e-5 .. 4e-5 with mirror padding and 2e-5 .. 1.1e-4 with zero padding at these sizes
    bar = 1e-4 if fm.prop_pad == "symmetric" else 2e-4
    assert e_d < bar and e_b < bar


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
    """GPU gradient vs GPU central differences along a direction supported on 64 random voxels.

    Single-voxel differences of an FP32 forward model sit below its rounding
    noise, so the 64 voxels are perturbed together (weights of the gradient's
    sign, step chosen for a 1e-3 relative change of L); per-voxel central
    differences are checked in FP64 against the oracle gradient
    (tests/test_oracle_model.py), which the GPU gradient matches to 1e-4."""
    fm, orc, data, delta, beta = setup
    dd, bb, yy = (torch.from_numpy(a).cuda() for a in (delta, beta, data))
    gd, gb = fm.grad(dd, bb, yy)
    L0 = fm.loss(dd, bb, yy).item()
    rng = np.random.default_rng(7)
    for which, g in ((0, gd), (1, gb)):
        idx = torch.from_numpy(rng.choice(delta.size, 64, replace=False)).cuda()
        v = torch.zeros(delta.size, dtype=torch.float64, device="cuda")
        gv = g.view(-1)[idx].double()
        v[idx] = torch.sign(gv) * torch.from_numpy(rng.uniform(0.5, 1.5, 64)).cuda()
        dirderiv = float(torch.dot(g.view(-1).double(), v))
        h = 5e-4 * L0 / abs(dirderiv)
        vols_p = [dd.clone(), bb.clone()]
        vols_m = [dd.clone(), bb.clone()]
        vols_p[which] = (vols_p[which].double() + h * v.view(delta.shape)).float()
        vols_m[which] = (vols_m[which].double() - h * v.view(delta.shape)).float()
        fd = (fm.loss(*vols_p, yy).item() - fm.loss(*vols_m, yy).item()) / (2 * h)
        err = abs(fd - dirderiv) / abs(dirderiv)
        print(f"FD check ({'delta' if which == 0 else 'beta'}): 64-voxel directional rel err {err:.2e}")
        assert err < 1e-3


def test_forward_token_reuse_and_staleness(setup):
    """A backward given its forward's token starts from the sinograms and wavefield that forward left and equals
    the backward that reruns the forward (token 0) bit for bit; a token made stale by a later forward, or
    used up by a backward, falls back to rerunning the forward."""
    fm, orc, data, delta, beta = setup
    dd, bb, yy = (torch.from_numpy(a).cuda() for a in (delta, beta, data))
    fresh = fm.grad(dd, bb, yy)
    _, tok = fm._forward(dd, bb, yy)
    assert tok != 0
    reuse = fm.grad(dd, bb, yy, token=tok)
    assert all(torch.equal(a, b) for a, b in zip(reuse, fresh))
    again = fm.grad(dd, bb, yy, token=tok)          # used up by the backward above
    assert all(torch.equal(a, b) for a, b in zip(again, fresh))
    _, tok1 = fm._forward(dd, bb, yy)
    fm._forward(2 * dd, bb, yy)                      # a later forward on other inputs
    stale = fm.grad(dd, bb, yy, token=tok1)
    assert all(torch.equal(a, b) for a, b in zip(stale, fresh))
