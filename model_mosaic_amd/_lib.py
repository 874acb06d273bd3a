"""ctypes binding of libmosaic_hip.so (the C ABI declared in include/mosaic_hip.h).

The library is loaded from this package directory only.  There is no fallback:
if it is missing or fails to load, every operator raises ``MosaicLibraryError``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmosaic_hip.so"

MM_OK = 0
MM_ERR_ARG = -1
MM_ERR_OOM = -2
MM_ERR_HIP = -3
MM_ERR_UNSUPPORTED = -4
MM_RADON_FP64_TILES = 1
MM_RADON_GATHER_ZP1 = 2
MM_RADON_KAHAN_TILES = 4
MM_RADON_LIST_TILES = 8
MM_RADON_GATHER_ROOT32 = 16
MM_RADON_TUNE_GATHER_XCD_GROUP = 1
MM_RADON_TUNE_SCATTER_XCD_GROUP = 2
MM_RADON_TUNE_COL_SPLIT = 3
MM_RADON_TUNE_GATHER_ROOT = 4
MM_RADON_TUNE_GATHER_ZP = 5
MM_MOSAIC_NO_GRAPH = 1
MM_MOSAIC_AUTOTUNE = 2
MM_MOSAIC_INFO_RADON_PLANS = 0
MM_MOSAIC_INFO_FRESNEL_PLANS = 1
MM_MOSAIC_INFO_CAPTURES = 2
MM_MOSAIC_INFO_TILES = 3
MM_MOSAIC_INFO_BRANCHES = 4
MM_MOSAIC_TILE_BRANCH = 100


def MM_MOSAIC_BRANCHES(b: int) -> int:
    return int(b) << 8
MM_RADON_NSTAGES = 8
RADON_STAGES = ("r_rows_rfft", "r_col_fft", "r_gather", "r_line_ifft",
                "rt_line_fft", "rt_scatter", "rt_col_ifft", "rt_rows_c2r")

# Every entry point of include/mosaic_hip.h: name -> (restype, argtypes)
_vp = C.c_void_p
_SIGS = {
    "mm_last_error": (C.c_char_p, []),
    "mm_version": (C.c_int, []),
    "mm_radon_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_uint]),
    "mm_radon_fwd": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "mm_radon_fwd_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "mm_radon_adj": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp]),
    "mm_radon_tune": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mm_radon_plan_check": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "mm_radon_get_tune": (C.c_int, [_vp, C.c_int]),
    "mm_radon_autotune": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    "mm_radon_profile": (C.c_int, [_vp, C.c_int]),
    "mm_radon_stage_ms": (C.c_int, [_vp, C.POINTER(C.c_float), C.c_int]),
    "mm_radon_workspace_bytes": (C.c_size_t, [_vp]),
    "mm_radon_destroy": (None, [_vp]),
    "mm_fresnel_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.c_int]),
    "mm_fresnel_fwd": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "mm_fresnel_adj": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "mm_fresnel_workspace_bytes": (C.c_size_t, [_vp]),
    "mm_fresnel_destroy": (None, [_vp]),
    "mm_paganin_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double]),
    "mm_paganin_retrieve": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "mm_paganin_workspace_bytes": (C.c_size_t, [_vp]),
    "mm_paganin_destroy": (None, [_vp]),
    "mm_fbp_filter": (C.c_int, [_vp, _vp, C.c_longlong, C.c_int, _vp]),
    "mm_fbp_filter_2n": (C.c_int, [_vp, _vp, _vp, _vp, C.c_longlong, C.c_int, _vp]),
    "mm_radon_adj_fbp": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp]),
    "mm_psi_from_proj": (C.c_int, [_vp, _vp, C.c_longlong, C.c_float, C.c_float, _vp]),
    "mm_abs2": (C.c_int, [_vp, _vp, C.c_longlong, _vp]),
    "mm_mosaic_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.c_float, C.c_float,
                                   C.c_int, C.c_uint]),
    "mm_mosaic_run": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _vp]),
    "mm_mosaic_info": (C.c_int, [_vp, C.c_int]),
    "mm_mosaic_tile_knob": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mm_mosaic_workspace_bytes": (C.c_size_t, [_vp]),
    "mm_mosaic_destroy": (None, [_vp]),
    "mm_plain_gather": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "mm_plain_scatter": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _vp]),
    "mm_model_create": (C.c_int, [C.POINTER(_vp), _vp, _vp, C.c_int, C.c_float]),
    "mm_model_fwd": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp, C.POINTER(C.c_ulonglong), _vp]),
    "mm_model_bwd": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_float, C.c_ulonglong, _vp, _vp, _vp]),
    "mm_model_workspace_bytes": (C.c_size_t, [_vp]),
    "mm_model_destroy": (None, [_vp]),
}

EXPORTED = tuple(_SIGS)


class MosaicLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or a call failed."""


_lib = None


def load(path: os.PathLike | str | None = None):
    """Loads (once) and returns the library handle with typed entry points."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise MosaicLibraryError(
            f"{p} is missing: build it with `python -m model_mosaic_amd.build` (hipcc, gfx950)")
    try:
        lib = C.CDLL(str(p))
    except OSError as e:
        raise MosaicLibraryError(f"cannot load {p}: {e}") from e
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    msg = load().mm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str):
    """Maps a status code to the reference's exception types."""
    if rc == MM_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == MM_ERR_ARG:
        raise ValueError(msg)
    if rc == MM_ERR_OOM:
        raise MemoryError(msg)
    if rc == MM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise MosaicLibraryError(msg)


def stream_handle(stream=None) -> int:
    """HIP stream handle of a torch stream (default: torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
