"""HIP Radon R / RT parity against the oracle and its golden fixtures (needs a GPU)."""
import os

import numpy as np
import pytest
import torch

from oracle import usfft

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def tomo_cls():
    from model_mosaic_amd.processing.tomo import TomoReal
    return TomoReal


def test_R_RT_golden_small(tomo_cls):
    g = gold("radon_small")
    nz = g["obj"].shape[0]
    t = tomo_cls(64, nz, g["theta"])
    r = t.R(torch.from_numpy(g["obj"]).cuda()).cpu().numpy()
    assert r.dtype == np.float32 and r.shape == g["R"].shape
    b = t.RT(torch.from_numpy(g["y"]).cuda()).cpu().numpy()
    c = t.RT(torch.from_numpy(g["yc"]).cuda()).cpu().numpy()
    print(f"small: R {rel(r, g['R']):.2e}  RT {rel(b, g['RT']):.2e}  RT complex {rel(c, g['RTc']):.2e}")
    assert rel(r, g["R"]) < 1e-5
    assert b.dtype == np.float32 and rel(b, g["RT"]) < 1e-5
    assert c.dtype == np.complex64 and rel(c, g["RTc"]) < 1e-5


def test_R_RT_golden_c1(tomo_cls):
    g = gold("radon_c1")
    t = tomo_cls(256, 2, g["theta"])
    r = t.R(torch.from_numpy(g["obj"]).cuda()).cpu().numpy()
    err64 = rel(r, g["R"])
    err_nudft = rel(r[:, :1], g["R_nudft"])
    print(f"C1 sizes: R rel-L2 vs FP64 USFFT {err64:.2e}, vs exact NUDFT {err_nudft:.2e}")
    assert err64 < 1e-5
    assert err_nudft < 2e-3
    b = t.RT(torch.from_numpy(g["R"]).cuda()).cpu().numpy()
    assert rel(b, g["RT"]) < 1e-5


def test_R_matches_fp32_reference_restatement(tomo_cls):
    n, nz, nt = 64, 3, 40
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    x = np.random.default_rng(11).standard_normal((nz, n, n)).astype(np.float32)
    r = tomo_cls(n, nz, th).R(torch.from_numpy(x).cuda()).cpu().numpy()
    assert rel(r, usfft.tomo_R(x, th, np.float32)) < 1e-5
    assert rel(r, usfft.tomo_R(x, th, np.float64)) < 1e-5


@pytest.mark.parametrize("n,nt,span,nz,zp", [(16, 7, np.pi, 2, 2), (32, 33, np.pi, 3, 2), (128, 64, 2 * np.pi, 2, 2),
                                             (512, 96, np.pi, 2, 2), (128, 64, 2 * np.pi, 3, 1), (512, 96, np.pi, 5, 1),
                                             (256, 80, np.pi, 11, 2)])
def test_sizes_vs_fp64(tomo_cls, n, nt, span, nz, zp):
    """Both gather variants (one or two z slices per patch element), odd and even nz."""
    th = np.linspace(0, span, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(n + nt)
    x = rng.standard_normal((nz, n, n)).astype(np.float32)
    y = rng.standard_normal((nt, nz, n)).astype(np.float32)
    t = tomo_cls(n, nz, th, gather_zp=zp)
    assert rel(t.R(torch.from_numpy(x).cuda()).cpu().numpy(), usfft.tomo_R(x, th, np.float64)) < 1e-5
    assert rel(t.RT(torch.from_numpy(y).cuda()).cpu().numpy(), usfft.tomo_RT(y, th, dtype=np.float64)) < 1e-5


@pytest.mark.parametrize("n,nt,nz", [(96, 37, 3), (192, 50, 2), (768, 129, 5), (1536, 64, 2),
                                     # long transforms: two-column / one-pair workgroups kept paired per XCD
                                     (3072, 24, 1), (4096, 16, 1)])
def test_sizes_3x2k_vs_fp64(tomo_cls, n, nt, nz):
    """Detector widths 3 * 2^k (the mosaic tiles' 1536 and 3072): radix-12 FFT passes, non-power-of-two
    Fourier grid wrap-around in the gather and scatter."""
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(n + nt)
    x = rng.standard_normal((nz, n, n)).astype(np.float32)
    y = rng.standard_normal((nt, nz, n)).astype(np.float32)
    t = tomo_cls(n, nz, th)
    r = t.R(torch.from_numpy(x).cuda()).cpu().numpy()
    b = t.RT(torch.from_numpy(y).cuda()).cpu().numpy()
    er, eb = rel(r, usfft.tomo_R(x, th, np.float64)), rel(b, usfft.tomo_RT(y, th, dtype=np.float64))
    print(f"n={n}: R {er:.2e}  RT {eb:.2e}")
    assert er < 1e-5 and eb < 1e-5


def test_random_angles_and_partial_nz(tomo_cls):
    n, nt = 64, 29
    th = np.sort(np.random.default_rng(5).uniform(0, 2 * np.pi, nt)).astype(np.float32)
    x = np.random.default_rng(6).standard_normal((3, n, n)).astype(np.float32)
    t = tomo_cls(n, 8, th)   # plan for 8 slices, call with 3
    assert rel(t.R(torch.from_numpy(x).cuda()).cpu().numpy(), usfft.tomo_R(x, th, np.float64)) < 1e-5
    # RT with 5 of 8 slices: the second z quad of the staging round is partial
    y = np.random.default_rng(7).standard_normal((nt, 5, n)).astype(np.float32)
    assert rel(t.RT(torch.from_numpy(y).cuda()).cpu().numpy(), usfft.tomo_RT(y, th, dtype=np.float64)) < 1e-5


@pytest.mark.parametrize("fp64_tiles,kahan", [(False, False), (True, False), (False, True)])
def test_adjoint_identity(tomo_cls, fp64_tiles, kahan):
    """<R x, y> = <x, RT y> (tests/test_large_variants.py:122-172), inner products in FP64.

    Random x, y make <Rx, y> cancel down to ~sqrt(N) of its terms, which
    magnifies FP32 rounding; so besides the reference's |lhs-rhs|/max(|lhs|,|rhs|)
    (at its own tolerance) the error is also bounded by the Cauchy-Schwarz scale
    ||Rx|| ||y|| and on the non-cancelling pair y = Rx, both at 1e-6."""
    n, nz, nt = 64, 8, 24
    th = np.linspace(0, 2 * np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.standard_normal((nz, n, n)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal((nt, nz, n)).astype(np.float32)).cuda()
    t = tomo_cls(n, nz, th, fp64_tiles=fp64_tiles, kahan_tiles=kahan)
    rx = t.R(x).double()
    lhs = torch.sum(rx * y.double()).item()
    rhs = torch.sum(x.double() * t.RT(y).double()).item()
    ref_err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    cs_err = abs(lhs - rhs) / (rx.norm().item() * y.double().norm().item())
    lhs2 = torch.sum(rx * rx).item()
    rhs2 = torch.sum(x.double() * t.RT(rx.float()).double()).item()
    self_err = abs(lhs2 - rhs2) / lhs2
    print(f"adjoint: ref-metric {ref_err:.2e}  cauchy-schwarz {cs_err:.2e}  y=Rx {self_err:.2e}")
    assert ref_err < 1e-4
    assert cs_err < 1e-6
    assert self_err < 1e-6


def test_R_deterministic(tomo_cls):
    n, nz, nt = 128, 2, 64
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    x = torch.randn((nz, n, n), device="cuda")
    t = tomo_cls(n, nz, th)
    assert torch.equal(t.R(x), t.R(x))


def test_shape_errors(tomo_cls):
    t = tomo_cls(32, 2, np.linspace(0, np.pi, 8, endpoint=False))
    with pytest.raises(ValueError):
        t.R(torch.zeros((3, 32, 32), device="cuda"))
    with pytest.raises(ValueError):
        t.R(torch.zeros((2, 32, 16), device="cuda"))
    with pytest.raises(ValueError):
        t.RT(torch.zeros((7, 2, 32), device="cuda"))


# ---- TomoLargeReal: tests/test_large_variants.py semantics ----------------------------------

def test_large_R_chunk_invariance_and_parity(tomo_cls):
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    n, nz, nt = 64, 16, 24
    th = np.linspace(0, 2 * np.pi, nt, endpoint=False).astype(np.float32)
    np.random.seed(0)
    obj = np.random.randn(nz, n, n).astype(np.float32)
    ref = tomo_cls(n, nz, th).R(torch.from_numpy(obj).cuda()).cpu().numpy()
    tl = TomoLargeReal(n, th, n)
    coarse = tl.R(obj, [n, nt, n]).copy()
    for chunks in [(32, 12, 32), (16, 6, 16), (8, 6, 8)]:
        fine = TomoLargeReal(n, th, chunks[2]).R(obj, list(chunks)).copy()
        assert np.max(np.abs(fine - coarse)) / np.max(np.abs(coarse)) < 1e-4
    assert np.max(np.abs(coarse - ref)) / np.max(np.abs(ref)) < 1e-4
    assert rel(coarse, usfft.tomo_large_R(obj, th, np.float64)) < 1e-5


def test_large_R_returns_view_until_next_call():
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    n, nt = 32, 8
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    tl = TomoLargeReal(n, th, n)
    a = tl.R(np.ones((2, n, n), np.float32), [n, nt, n])
    keep = a.copy()
    tl.R(np.zeros((2, n, n), np.float32), [n, nt, n])
    assert np.all(a == 0) and np.any(keep != 0)   # same buffer, overwritten (tomo_large.py:545-553)


@pytest.mark.parametrize("cplx", [True, False])
def test_large_RT_parity(tomo_cls, cplx):
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    n, nz, nt = 64, 8, 24
    th = np.linspace(0, 2 * np.pi, nt, endpoint=False).astype(np.float32)
    np.random.seed(0)
    if cplx:
        sino = (np.random.randn(nt, nz, n) + 1j * np.random.randn(nt, nz, n)).astype(np.complex64)
    else:
        sino = np.random.randn(nt, nz, n).astype(np.float32)
    ref = tomo_cls(n, nz, th).RT(torch.from_numpy(sino).cuda()).cpu().numpy()
    for chunks in [(n, nt, 2 * n), (32, 12, 32), (16, 6, 16), (8, 6, 8)]:
        got = np.asarray(TomoLargeReal(n, th, chunks[2]).RT(sino, list(chunks))).copy()
        assert got.dtype == sino.dtype
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-4
    assert rel(got, usfft.tomo_large_RT(sino, th, np.float64)) < 1e-5


def test_large_adjoint_and_validation():
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    n, nz, nt = 64, 8, 24
    th = np.linspace(0, 2 * np.pi, nt, endpoint=False).astype(np.float32)
    np.random.seed(0)
    x = np.random.randn(nz, n, n).astype(np.float32)
    y = np.random.randn(nt, nz, n).astype(np.float32)
    tl = TomoLargeReal(n, th, n)
    rx = np.asarray(tl.R(x, [n, nt, n])).copy()
    rty = np.asarray(tl.RT(y, [n, nt, n])).copy()
    lhs, rhs = float(np.sum(rx.astype(np.float64) * y)), float(np.sum(x.astype(np.float64) * rty))
    assert abs(lhs - rhs) / max(abs(lhs), abs(rhs)) < 1e-6
    with pytest.raises(AssertionError):
        tl.R(x, [24, nt, n])


@pytest.mark.parametrize("n,nt", [(64, 24), (128, 50)])
def test_plain_reference_structured_path(n, nt):
    """The benchmark's plain one-slice comparator computes the same operator (FP64 oracle, 1e-5)."""
    from model_mosaic_amd.plain_path import PlainTomo
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((1, n, n)).astype(np.float32)
    y = rng.standard_normal((nt, 1, n)).astype(np.float32)
    p = PlainTomo(n, th)
    r = p.R(torch.from_numpy(x[0]).cuda()).cpu().numpy()
    print(f"plain R rel-L2 vs FP64 {rel(r, usfft.tomo_R(x, th, np.float64)[:, 0]):.2e}")
    b = p.RT(torch.from_numpy(y[:, 0]).cuda()).cpu().numpy()
    assert rel(r, usfft.tomo_R(x, th, np.float64)[:, 0]) < 1e-5
    assert rel(b, usfft.tomo_RT(y, th, dtype=np.float64)[0]) < 1e-5


@pytest.mark.parametrize("cplx", [False, True])
def test_large_streams_z_slabs_bitwise(tomo_cls, cplx):
    """A GPU budget of two slices makes TomoLargeReal stream 7 slices as 4 z slabs through the three-stream
    pipeline (uploads and downloads overlapping the transforms): the result equals the one-batch TomoReal
    bit for bit (slices are independent), for R and for RT of real and complex sinograms."""
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    n, nz, nt = 64, 7, 24
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(9)
    probe = TomoLargeReal(n, th, n)
    budget = 2 * probe._slice_bytes(cplx) + 1
    tl = TomoLargeReal(n, th, n, gpu_budget_bytes=budget)
    assert tl._zbatch(nz, cplx) == 2
    t = tomo_cls(n, nz, th)
    if cplx:
        sino = (rng.standard_normal((nt, nz, n)) + 1j * rng.standard_normal((nt, nz, n))).astype(np.complex64)
    else:
        sino = rng.standard_normal((nt, nz, n)).astype(np.float32)
        obj = rng.standard_normal((nz, n, n)).astype(np.float32)
        assert np.array_equal(tl.R(obj, [n, nt, n]), t.R(torch.from_numpy(obj).cuda()).cpu().numpy())
    got = tl.RT(sino, [n, nt, n])
    assert got.dtype == sino.dtype
    assert np.array_equal(got, t.RT(torch.from_numpy(sino).cuda()).cpu().numpy())


@pytest.mark.parametrize("n,nt,nz", [(256, 384, 2), (2048, 256, 1), (64, 24, 17), (96, 200, 33)])
def test_scatter_accumulators_vs_fp64(tomo_cls, n, nt, nz):
    """RT with the scatter as dense patch products on the matrix cores (default), as FP32 per-point lists on the
    vector ALUs, as compensated FP32 (Kahan) lists and as FP64 lists, each vs the FP64 oracle: all within the 1e-5
    bar, the compensated and FP64 sums no worse than plain FP32 lists (beyond rounding noise), the matrix-core
    result within FP32 rounding of the list result (nz = 17, 33: partial staging rounds of 16 slices; plans with
    nz < 16 take the lists by default, so there "mfma" and "lists" are the same path)."""
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    y = np.random.default_rng(n + 1).standard_normal((nt, nz, n)).astype(np.float32)
    ref = usfft.tomo_RT(y, th, dtype=np.float64)
    errs, outs = {}, {}
    for name, kw in (("mfma", {}), ("lists", {"list_tiles": True}), ("kahan", {"kahan_tiles": True}),
                     ("fp64", {"fp64_tiles": True})):
        outs[name] = tomo_cls(n, nz, th, **kw).RT(torch.from_numpy(y).cuda()).cpu().numpy()
        errs[name] = rel(outs[name], ref)
    d = rel(outs["mfma"], outs["lists"])
    print(f"RT n={n} ntheta={nt} nz={nz}: rel-L2 vs FP64 " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items())
          + f"; matrix cores vs lists {d:.2e}")
    assert max(errs.values()) < 1e-5
    assert errs["kahan"] <= 1.05 * errs["lists"] and errs["fp64"] <= 1.05 * errs["lists"]
    assert d < 1e-6
    for bad in ({"fp64_tiles": True, "kahan_tiles": True}, {"list_tiles": True, "kahan_tiles": True},
                {"list_tiles": True, "fp64_tiles": True}):
        with pytest.raises(ValueError):
            tomo_cls(64, 1, th[:8], **bad)


@pytest.mark.parametrize("n,nt,nz", [(256, 180, 3), (2048, 64, 2)])
def test_gather_shape_knobs_bitwise(tomo_cls, n, nt, nz):
    """Rebuilding the gather tables for root tile 64 / 32 and one / two slices per patch element (the knobs the
    autotuner sweeps) leaves R bitwise unchanged, keeps the plan's workspace byte count equal to a plan created
    with the same shape, and rejects values outside {32, 64} / {1, 2}; the autotuner keeps R's bits too."""
    import ctypes as C
    from model_mosaic_amd import _lib
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    x = torch.from_numpy(np.random.default_rng(n).standard_normal((nz, n, n)).astype(np.float32)).cuda()
    t = tomo_cls(n, nz, th)
    ref = t.R(x).clone()
    lib = t._lib
    for root in (32, 64):
        for zp in (1, 2):
            _lib.check(lib.mm_radon_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ROOT, root), "tune")
            _lib.check(lib.mm_radon_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ZP, zp), "tune")
            assert lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ROOT) == root
            assert lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ZP) == zp
            assert torch.equal(t.R(x), ref), (root, zp)
            fresh = tomo_cls(n, nz, th, gather_root=root, gather_zp=zp)
            assert t.workspace_bytes() == fresh.workspace_bytes(), (root, zp)
            fresh.free()
    for knob, bad in ((_lib.MM_RADON_TUNE_GATHER_ROOT, 16), (_lib.MM_RADON_TUNE_GATHER_ROOT, 128),
                      (_lib.MM_RADON_TUNE_GATHER_ZP, 0), (_lib.MM_RADON_TUNE_GATHER_ZP, 4)):
        assert lib.mm_radon_tune(t._plan, knob, bad) == _lib.MM_ERR_ARG
    sino = torch.empty_like(ref)
    back = torch.empty_like(x)
    _lib.check(lib.mm_radon_autotune(t._plan, C.c_void_p(x.data_ptr()), C.c_void_p(sino.data_ptr()),
                                     C.c_void_p(back.data_ptr()), nz, _lib.stream_handle()), "autotune")
    assert lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ROOT) in (32, 64)
    assert lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_GATHER_ZP) in (1, 2)
    assert torch.equal(t.R(x), ref)
    t.free()


@pytest.mark.parametrize("n,nt,nz", [(1024, 96, 2), (1536, 64, 2), (2048, 64, 2)])
def test_column_split_knob_bitwise(tomo_cls, n, nt, nz):
    """Half as many columns per column-FFT workgroup (MM_RADON_TUNE_COL_SPLIT = 1, the default at n = 1024 and
    1536) and the full-size workgroups (0) give R and RT bit for bit the same: the per-column transforms are
    identical, only their packing into workgroups changes."""
    from model_mosaic_amd import _lib
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(n + nt)
    x = torch.from_numpy(rng.standard_normal((nz, n, n)).astype(np.float32)).cuda()
    y = torch.from_numpy(rng.standard_normal((nt, nz, n)).astype(np.float32)).cuda()
    t = tomo_cls(n, nz, th)
    assert t._lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_COL_SPLIT) == (1 if n in (1024, 1536) else 0)
    outs = []
    for knob in (0, 1):
        _lib.check(t._lib.mm_radon_tune(t._plan, _lib.MM_RADON_TUNE_COL_SPLIT, knob), "tune")
        outs.append((t.R(x).clone(), t.RT(y).clone()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    t.free()


@pytest.mark.parametrize("n,nt,nz", [(256, 384, 16), (2048, 256, 16)])
def test_scatter_xcd_group_bitwise(tomo_cls, n, nt, nz):
    """The scatter's XCD group size (MM_RADON_TUNE_SCATTER_XCD_GROUP: 64 by default from n = 2048 on, 32 below)
    only reorders the work items over the chip: every tile and partial slot has one owner and scatter_reduce
    sums the slots in a fixed order, so RT is bit for bit the same for every group size."""
    from model_mosaic_amd import _lib
    th = np.linspace(0, np.pi, nt, endpoint=False).astype(np.float32)
    rng = np.random.default_rng(7 * n + nt)
    y = torch.from_numpy(rng.standard_normal((nt, nz, n)).astype(np.float32)).cuda()
    t = tomo_cls(n, nz, th)
    assert t._lib.mm_radon_get_tune(t._plan, _lib.MM_RADON_TUNE_SCATTER_XCD_GROUP) == (64 if n >= 2048 else 32)
    outs = []
    for g in (1, 4, 32, 64):
        _lib.check(t._lib.mm_radon_tune(t._plan, _lib.MM_RADON_TUNE_SCATTER_XCD_GROUP, g), "tune")
        outs.append(t.RT(y).clone())
    assert all(torch.equal(outs[0], o) for o in outs[1:])
    t.free()
