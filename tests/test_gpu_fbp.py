"""HIP FBP filter parity against the oracle (needs a GPU)."""
import numpy as np
import pytest
import torch

from oracle import fbp

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("method", ["2n", "4n"])
@pytest.mark.parametrize("n,rows,name", [(64, 7, "shepp"), (96, 10, "ramp"), (256, 33, "parzen"),
                                         (2048, 9, "shepp"), (1536, 4, "hann"), (16, 3, "cosine")])
def test_filter_vs_fp64(n, rows, name, method):
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    f = FBPFilter(n, method=method)
    w = f.calc_filter(name)
    x = np.random.default_rng(n + rows).standard_normal((rows, n)).astype(np.float32)
    d = torch.from_numpy(x).cuda()
    out = f.filter(d, w)
    assert out.data_ptr() == d.data_ptr()          # in place, like the reference
    ref = fbp.filter(x, w.cpu().numpy().astype(np.float64))
    err = rel(d.cpu().numpy(), ref)
    print(f"fbp {method} {name} n={n} rows={rows}: rel-L2 vs FP64 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("method", ["2n", "4n"])
def test_none_filter_closed_form_and_3d_sinogram(method):
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    n = 128
    ne = 4 * n
    f = FBPFilter(n, method=method)
    x = np.random.default_rng(3).standard_normal((5, 3, n)).astype(np.float32)   # (ntheta, nz, n)
    d = torch.from_numpy(x).cuda()
    f.filter(d, f.calc_filter("none"))
    ref = 3 * n * ne * x.astype(np.float64) + 9 * n * fbp.edge_pad(x.astype(np.float64), ne).sum(-1, keepdims=True)
    assert rel(d.cpu().numpy(), ref) < 1e-6


def test_filter_host_and_errors():
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    n = 64
    f = FBPFilter(n, batch_chunk=4)
    w = f.calc_filter("ramp")
    x = np.random.default_rng(4).standard_normal((6, 2, n)).astype(np.float32)
    h = x.copy()
    f.filter_host(h, w)
    assert rel(h, fbp.filter(x, w.cpu().numpy().astype(np.float64))) < 1e-5
    with pytest.raises(ValueError):
        f.filter(torch.zeros((3, 32), device="cuda"), w)
    with pytest.raises(ValueError):
        f.filter(torch.zeros((4, 2 * n), device="cuda")[:, :n], w)   # not contiguous
    with pytest.raises(ValueError):
        f.calc_filter("gauss")


def test_weights_changed_in_place_rebuild_the_2n_tables():
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    n = 64
    f = FBPFilter(n)
    w = f.calc_filter("ramp").clone()
    x = np.random.default_rng(5).standard_normal((4, n)).astype(np.float32)
    f.filter(torch.from_numpy(x).cuda(), w)
    w.mul_(0.5)
    d = torch.from_numpy(x).cuda()
    f.filter(d, w)
    assert rel(d.cpu().numpy(), fbp.filter(x, w.cpu().numpy().astype(np.float64))) < 1e-5


@pytest.mark.parametrize("n,nt,nz", [(64, 24, 3), (96, 40, 2), (256, 100, 4), (2048, 64, 2),
                                     (384, 40, 2)])  # 384: filter and line transforms differ in points per thread
def test_RT_filtered_equals_filter_then_RT(n, nt, nz):
    """TomoReal.RT_filtered (filter fused into RT's line pass) = FBPFilter.filter then TomoReal.RT to FP32
    rounding (same transforms; the compiler may contract the fused kernel's arithmetic differently), and
    both = the FP64 oracle chain; the input is untouched."""
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    from model_mosaic_amd.processing.tomo import TomoReal

    from oracle import usfft
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    t = TomoReal(n, nz, th)
    f = FBPFilter(n)
    w = f.calc_filter("shepp")
    y = np.random.default_rng(n).standard_normal((nt, nz, n)).astype(np.float32)
    d = torch.from_numpy(y).cuda()
    fused = t.RT_filtered(d, f, w).cpu().numpy()
    assert np.array_equal(d.cpu().numpy(), y)
    two = t.RT(f.filter(d.clone(), w)).cpu().numpy()
    e2 = rel(fused, two)
    print(f"RT_filtered n={n}: rel-L2 vs filter + RT {e2:.2e}")
    assert e2 < 1e-6
    ref = usfft.tomo_RT(fbp.filter(y, w.cpu().numpy().astype(np.float64)), th, dtype=np.float64)
    err = rel(fused, ref)
    print(f"RT_filtered n={n}: rel-L2 vs FP64 {err:.2e}")
    assert err < 1e-5


def test_weight_tables_follow_the_values_not_the_address():
    """Numpy weights are converted afresh on every call, and a freed tensor's address is reused by the
    next one: each call must still filter with its own weights (ramp, then shepp, then ramp modified in
    place), both for FBPFilter.filter and TomoReal.RT_filtered."""
    from model_mosaic_amd.processing.fbp_filter import FBPFilter, calc_filter_host
    from model_mosaic_amd.processing.tomo import TomoReal

    from oracle import usfft
    n, nt, nz = 64, 24, 2
    f = FBPFilter(n)
    x = np.random.default_rng(21).standard_normal((nt, nz, n)).astype(np.float32)
    weights = [calc_filter_host(n, "ramp"), calc_filter_host(n, "shepp")]
    for w in weights + weights[::-1]:
        d = torch.from_numpy(x).cuda()
        f.filter(d, w)
        assert rel(d.cpu().numpy(), fbp.filter(x, w.astype(np.float64))) < 1e-5
    wt = torch.from_numpy(weights[0].copy()).cuda()
    d = torch.from_numpy(x).cuda()
    f.filter(d, wt)
    wt.mul_(0.5)                                   # same tensor, changed in place
    d = torch.from_numpy(x).cuda()
    f.filter(d, wt)
    assert rel(d.cpu().numpy(), fbp.filter(x, 0.5 * weights[0].astype(np.float64))) < 1e-5
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    t = TomoReal(n, nz, th)
    for w in weights:
        got = t.RT_filtered(torch.from_numpy(x).cuda(), f, w).cpu().numpy()
        ref = usfft.tomo_RT(fbp.filter(x, w.astype(np.float64)), th, dtype=np.float64)
        assert rel(got, ref) < 1e-5
    assert len(f._tables) <= f._TABLES_KEPT


def test_out_buffers_are_checked():
    """A caller's out= of the wrong size or dtype is rejected before any kernel writes into it."""
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.processing.tomo import TomoReal
    n, nt, nz = 32, 8, 2
    t = TomoReal(n, nz, np.linspace(0, np.pi, nt, endpoint=False))
    x = torch.zeros((nz, n, n), device="cuda")
    with pytest.raises(ValueError):
        t.R(x, out=torch.empty((nt, nz, n - 1), device="cuda"))
    with pytest.raises(ValueError):
        t.RT(torch.zeros((nt, nz, n), device="cuda"), out=torch.empty((nz, n, n), dtype=torch.complex64, device="cuda"))
    p = Propagation(n, 16, 2, 1, 4.13e-11, 1.4e-6, [0.1])
    psi = torch.ones((16, n), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        p.D(psi, 0, out=torch.empty((16, n), dtype=torch.float32, device="cuda"))
    o = torch.empty((16, n), dtype=torch.complex64, device="cuda")
    res = p.D(psi, 0, out=o)
    assert res is o and res.shape == (16, n)       # a 2-D out comes back whole, not as one row
    assert torch.allclose(res, p.D(psi, 0))
