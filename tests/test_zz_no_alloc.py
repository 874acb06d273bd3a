"""mm_radon_fwd_noalloc (R with its half spectrum and representatives in the library's static device scratch)
against mm_radon_fwd (the plan's own workspace), called through the C ABI (needs a GPU).

The two entry points run the same kernels in the same order, so the sinograms must agree bit for bit: on two
calls in a row, the second with fewer slices (the scratch keeps the first call's contents, which the second
call must overwrite where it reads), and at a second geometry whose scratch offsets differ.  Oversized calls
are rejected before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from model_mosaic_amd import _lib

pytestmark = pytest.mark.gpu


def make_plan(lib, n, nz_max, theta):
    plan = C.c_void_p()
    th = np.ascontiguousarray(np.asarray(theta, dtype=np.float32).astype(np.float64))
    _lib.check(lib.mm_radon_create(C.byref(plan), n, nz_max, len(th), th.ctypes.data_as(C.POINTER(C.c_double)), 0),
               "mm_radon_create")
    return plan


def run(lib, fn, plan, x, ntheta):
    nz, n = x.shape[0], x.shape[1]
    out = torch.full((ntheta, nz, n), float("nan"), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(fn(plan, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), nz, C.c_void_p(stream)), fn.__name__)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("n,ntheta,nz_first,nz_second", [(256, 384, 8, 3), (192, 100, 5, 2)])
def test_noalloc_equals_fwd_on_two_calls(n, ntheta, nz_first, nz_second):
    lib = _lib.load()
    torch.cuda.set_device(0)
    th = np.linspace(0, np.pi, ntheta, endpoint=False).astype(np.float32)
    plan = make_plan(lib, n, nz_first, th)
    try:
        g = torch.Generator(device="cuda").manual_seed(n)
        for nz in (nz_first, nz_second):
            print(f"noalloc: n={n} ntheta={ntheta} nz={nz}", flush=True)
            x = torch.randn((nz, n, n), generator=g, device="cuda")
            ref = run(lib, lib.mm_radon_fwd, plan, x, ntheta)
            got = run(lib, lib.mm_radon_fwd_noalloc, plan, x, ntheta)
            assert torch.isfinite(ref).all()
            assert torch.equal(got, ref), f"nz={nz}"
    finally:
        torch.cuda.synchronize()
        lib.mm_radon_destroy(plan)


def test_noalloc_rejects_calls_beyond_the_scratch():
    """n = 2048 with 2048 angles fits one slice in the 128 MiB scratch; two slices are refused with an argument
    error (host-side check, nothing is launched)."""
    lib = _lib.load()
    torch.cuda.set_device(0)
    th = np.linspace(0, np.pi, 2048, endpoint=False).astype(np.float32)
    plan = make_plan(lib, 2048, 2, th)
    try:
        x = torch.zeros((2, 2048, 2048), device="cuda")
        out = torch.empty((2048, 2, 2048), device="cuda")
        rc = lib.mm_radon_fwd_noalloc(plan, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), 2, None)
        assert rc == _lib.MM_ERR_ARG
        assert "bytes of scratch" in _lib.last_error()
        one = run(lib, lib.mm_radon_fwd_noalloc, plan, x[:1].contiguous(), 2048)
        assert torch.equal(one, torch.zeros_like(one))
    finally:
        torch.cuda.synchronize()
        lib.mm_radon_destroy(plan)
