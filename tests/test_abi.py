"""The C-ABI library loads and exports every entry point declared in include/mosaic_hip.h.

No compute call is made here: only entry points that return before touching a
device (argument validation, version, error string)."""
import ctypes as C
import os
import re

import pytest

from model_mosaic_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mosaic_hip.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", txt))
    return names


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        from model_mosaic_amd import build
        build.build()
    return _lib.load()


def test_header_matches_binding_table():
    assert header_functions() == set(_lib.EXPORTED)


def test_library_exports_every_symbol(lib):
    for name in header_functions():
        assert hasattr(lib, name), name


def test_version_and_error_string(lib):
    assert lib.mm_version() >= 100
    assert isinstance(_lib.last_error(), str)


def test_null_plan_is_an_argument_error(lib):
    rc = lib.mm_radon_fwd(None, None, None, 1, None)
    assert rc == _lib.MM_ERR_ARG
    assert "plan is NULL" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "mm_radon_fwd")
    assert lib.mm_fresnel_fwd(None, None, None, 1, 0, None) == _lib.MM_ERR_ARG
    assert lib.mm_model_fwd(None, None, None, None, 0, None, None, None) == _lib.MM_ERR_ARG
    assert lib.mm_model_bwd(None, None, None, None, 0, 1.0, 0, None, None, None) == _lib.MM_ERR_ARG


def test_bad_sizes_rejected_before_device_work(lib):
    p = C.c_void_p()
    th = (C.c_double * 4)(0, 0.1, 0.2, 0.3)
    for n in (100, 48, 160, 49152, 65536):   # not 2^k / 3 * 2^k, 3 * 2^k below 96, 5 * 2^k, n above the largest widths
        rc = lib.mm_radon_create(C.byref(p), n, 1, 4, th, 0)
        assert rc == _lib.MM_ERR_ARG and not p, n
    for n, nz in ((64, 40), (20, 64), (64, 12)):   # 5 * 2^3, 5 * 2^2, 3 * 2^2 below 24
        rc = lib.mm_fresnel_create(C.byref(p), n, nz, 1, 1e-10, 1e-6, (C.c_double * 1)(0.1), 0)
        assert rc == _lib.MM_ERR_ARG and not p, (n, nz)
    # a side above 8192 runs only on the frame-length DCT form: symmetric padding, both sides DCT lengths
    for n, nz, pad in ((16384, 64, 1), (16384, 8, 0), (24, 12288, 0)):
        rc = lib.mm_fresnel_create(C.byref(p), n, nz, 1, 1e-10, 1e-6, (C.c_double * 1)(0.1), pad)
        assert rc == _lib.MM_ERR_ARG and not p, (n, nz, pad)
    rc = lib.mm_psi_from_proj(None, None, -1, 1.0, 1.0, None)
    assert rc == _lib.MM_ERR_ARG


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.MosaicLibraryError):
        _lib.load(tmp_path / "nope.so")


def _header_ranges(entry, var):
    """The '2^k in [a, b] or 3 * 2^k in [c, d]' limits the header states for `var` of `entry`."""
    txt = open(HEADER).read()
    seg = txt[txt.index(f" * {entry}:"):]
    seg = seg[:seg.index("*/")]
    seg = " ".join(re.sub(r"\n\s*\*", " ", seg).split())
    m = re.search(re.escape(var) + r"[^\[]*?2\^k in \[(\d+), (\d+)\] or 3 \* 2\^k in \[(\d+), (\d+)\]", seg)
    assert m, f"no size range for {var} in the {entry} comment"
    return tuple(int(v) for v in m.groups())


def test_header_size_limits_match_the_library(lib):
    """The Radon and Fresnel size limits written in the header are the ones the library enforces: the ends
    of each range are accepted (host-only plan check for Radon) and the next sizes outside are rejected."""
    a, b, c, d = _header_ranges("mm_radon_create", "n")
    th = (C.c_double * 2)(0.0, 0.3)
    for n in (a, b, c, d):
        assert lib.mm_radon_plan_check(n, 2, th) == _lib.MM_OK, (n, _lib.last_error())
    for n in (a // 2, 2 * b, c // 2, 2 * d):
        assert lib.mm_radon_plan_check(n, 2, th) == _lib.MM_ERR_ARG, n
    p = C.c_void_p()
    za, zb, zc, zd = _header_ranges("mm_fresnel_create", "nz ")
    fa, fb, fc, fd = _header_ranges("mm_fresnel_create", ", n ")
    assert (za, zb, zc, zd) == (8, 16384, 24, 12288) and (fa, fb, fc, fd) == (8, 32768, 24, 24576)
    dist = (C.c_double * 1)(0.1)
    for n in (fa // 2, 2 * fb, fc // 2, 2 * fd):
        assert lib.mm_fresnel_create(C.byref(p), n, 16, 1, 1e-10, 1e-6, dist, 0) == _lib.MM_ERR_ARG and not p, n
    for nz in (za // 2, 2 * zb, zc // 2, 2 * zd):
        assert lib.mm_fresnel_create(C.byref(p), 16, nz, 1, 1e-10, 1e-6, dist, 0) == _lib.MM_ERR_ARG and not p, nz
