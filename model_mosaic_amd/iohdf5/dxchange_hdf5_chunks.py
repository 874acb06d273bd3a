"""dxchange HDF5 files of the pipeline, read and written through libmosaic_h5.so (HDF5 C library).

Mirrors iohdf5/dxchange_hdf5_chunks.py: ``tomo_info`` (:99), ``tomo_initx`` (:223; its banking plan
:182), ``tomo_writex`` (:332), ``read_projs_vchunkx`` (:595), ``read_slices_vchunkx`` (:636) and
``write_vchunkx`` (:681), with the same file layout: a master file whose ``/exchange/data``
(nproj, ny, nx) float32 is a virtual dataset over bank files ``<base>/<base>_data_%06d<ext>`` split
along the projection axis (``stype='proj'``) or the slice axis (``stype='slice'``), plus
``/exchange/theta`` in degrees (step2_radon.py:107,137).

Differences of form: the ``shm`` arguments take any writable C-contiguous float32 buffer (a
``multiprocessing.shared_memory.SharedMemory``, a numpy array or a CPU torch tensor, e.g. page-locked);
reads and writes run in the calling process on the HDF5 library (``ntasks`` is accepted and unused: the
library reads a virtual box from all its bank files in one call).  ``slabs_to_device`` streams slabs of
a file to the GPU through page-locked buffers, the next slab read while the previous one uploads.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent.parent / "libmosaic_h5.so"
STYPE = {"proj": 0, "slice": 1}


class _Info(C.Structure):
    _fields_ = [("shape", C.c_longlong * 3), ("is_virtual", C.c_int), ("chunks", C.c_longlong * 3),
                ("vchunks", C.c_longlong * 3), ("nbanks_per_svchunk", C.c_int), ("stype", C.c_int)]


_lib = None


def load():
    """The HDF5 layer's library (built next to the package by ``model_mosaic_amd.build``)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from model_mosaic_amd._lib import MosaicLibraryError
            raise MosaicLibraryError(f"{LIB_PATH} is missing: build it with `python -m model_mosaic_amd.build` "
                                     "(needs the HDF5 C library)")
        lib = C.CDLL(str(LIB_PATH))
        ll3 = C.POINTER(C.c_longlong)
        sigs = {
            "mm_h5_last_error": (C.c_char_p, []),
            "mm_h5_info_get": (C.c_int, [C.c_char_p, C.POINTER(_Info)]),
            "mm_h5_read": (C.c_int, [C.c_char_p, ll3, ll3, C.c_void_p]),
            "mm_h5_create": (C.c_int, [C.c_char_p, ll3, ll3, C.c_int, C.c_int, ll3]),
            "mm_h5_write": (C.c_int, [C.c_char_p, ll3, ll3, C.c_void_p]),
            "mm_h5_write_theta": (C.c_int, [C.c_char_p, C.c_void_p, C.c_longlong]),
            "mm_h5_read_theta": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_longlong)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = load().mm_h5_last_error().decode()
        raise (ValueError if rc == -1 else OSError)(f"{what}: {msg}")


def _ll3(v):
    return (C.c_longlong * 3)(*[int(x) for x in v])


def _buffer(shm, shape):
    """A float32 ndarray of ``shape`` over the caller's buffer (SharedMemory, ndarray or CPU tensor)."""
    if hasattr(shm, "buf"):
        return np.ndarray(shape=shape, dtype=np.float32, buffer=shm.buf)
    if hasattr(shm, "numpy") and not isinstance(shm, np.ndarray):
        shm = shm.numpy()
    a = np.asarray(shm)
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError("buffer must be C-contiguous float32")
    return a.reshape(-1)[: int(np.prod(shape))].reshape(shape)


def tomo_info(filename) -> dict:
    """Shape, layout and writer attributes of ``/exchange/data`` (iohdf5/dxchange_hdf5_chunks.py:99)."""
    info = _Info()
    _check(load().mm_h5_info_get(os.fsencode(filename), C.byref(info)), "tomo_info")
    shape = tuple(int(v) for v in info.shape)
    return {"nproj": shape[0], "ny": shape[1], "nx": shape[2], "dtype": np.dtype(np.float32),
            "is_virtual": bool(info.is_virtual), "nbanks_per_svchunk": int(info.nbanks_per_svchunk),
            "vchunks": tuple(int(v) for v in info.vchunks), "shape": shape,
            "stype": "proj" if info.stype == 0 else "slice",
            "chunks": tuple(int(v) for v in info.chunks) if info.chunks[0] else None}


def banking_plan(filename, shape, vchunks, nbanks, stype):
    """Bank file paths and item counts of iohdf5/dxchange_hdf5_chunks.py:182-221."""
    axis = 0 if stype == "proj" else 1
    base, ext = os.path.splitext(os.path.basename(filename))
    fmt = os.path.join(os.path.dirname(filename), base, base) + "_data_%06d" + ext
    per = (vchunks[axis] + nbanks - 1) // nbanks
    paths, sizes = [], []
    for c in range((shape[axis] + vchunks[axis] - 1) // vchunks[axis]):
        for b in range(nbanks):
            off = c * vchunks[axis]
            lo, hi = off + b * per, min(off + (b + 1) * per, off + vchunks[axis], shape[axis])
            paths.append(fmt % (c * nbanks + b))
            sizes.append(max(0, hi - lo))
    return paths, sizes


def tomo_initx(filename, shape, dtype=np.float32, vchunks=None, mode="a", stype="proj", nbanks=1, chunks=None):
    """Creates the bank files and the virtual master (iohdf5/dxchange_hdf5_chunks.py:223-330); returns the
    context (bank paths and sizes) that ``tomo_writex`` takes."""
    if np.dtype(dtype) != np.float32:
        raise ValueError("tomo_initx: the containers hold float32")
    stype = "proj" if stype.lower() in ("proj", "projs") else "slice"
    vchunks = tuple(int(v) for v in (vchunks if vchunks is not None else shape))
    if os.path.isfile(filename) and mode == "a":
        try:
            tomo_info(filename)
            raise Exception("Virtual dataset exists already.")
        except OSError:
            pass
    ch = None if chunks is None else _ll3(chunks)
    _check(load().mm_h5_create(os.fsencode(filename), _ll3(shape), _ll3(vchunks), STYPE[stype], int(nbanks), ch),
           "tomo_initx")
    paths, sizes = banking_plan(filename, tuple(shape), vchunks, int(nbanks), stype)
    return {"banks_filename_path": paths, "banks_size": sizes}


def write_box(filename, start, data):
    """Writes the float32 array ``data`` at ``start`` of ``/exchange/data`` (into the bank files it covers)."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("write_box: data must be 3-D")
    _check(load().mm_h5_write(os.fsencode(filename), _ll3(start), _ll3(a.shape), a.ctypes.data), "write_box")


def read_box(filename, start, count, out=None):
    """The box (start, count) of ``/exchange/data`` into ``out`` (any float32 buffer; default a new array)."""
    count = tuple(int(c) for c in count)
    dst = np.empty(count, np.float32) if out is None else _buffer(out, count)
    _check(load().mm_h5_read(os.fsencode(filename), _ll3(start), _ll3(count), dst.ctypes.data), "read_box")
    return dst


def tomo_writex(filename, data, shm=None, ivchunk=(0, 0, 0), ctx=None):
    """Writes vchunk ``ivchunk`` (iohdf5/dxchange_hdf5_chunks.py:332): the box ivchunk * vchunks, clipped to
    the dataset, from ``data``."""
    info = tomo_info(filename)
    vc, shape = info["vchunks"], info["shape"]
    start = [int(i) * v for i, v in zip(ivchunk, vc)]
    count = [min(v, s - a) for v, s, a in zip(vc, shape, start)]
    a = np.asarray(data, dtype=np.float32)
    write_box(filename, start, a[: count[0], : count[1], : count[2]])


def read_projs_vchunkx(filename, shm, ntasks, vchunksx, ivchunkx, shard_axis=None):
    """Projection slab ivchunkx[0] of size vchunksx (iohdf5/dxchange_hdf5_chunks.py:595) into ``shm``;
    returns the vchunksx-shaped view."""
    info = tomo_info(filename)
    data = _buffer(shm, tuple(vchunksx))
    t0 = vchunksx[0] * ivchunkx[0]
    t1 = min(t0 + vchunksx[0], info["shape"][0])
    nz, nx = min(vchunksx[1], info["shape"][1]), min(vchunksx[2], info["shape"][2])
    if t1 > t0:
        tmp = read_box(filename, (t0, 0, 0), (t1 - t0, nz, nx))
        data[: t1 - t0, :nz, :nx] = tmp
    return data


def read_slices_vchunkx(filename, shm, ntasks, vchunksx, ivchunkx, shard_axis=None):
    """Slice slab ivchunkx[1] of size vchunksx (iohdf5/dxchange_hdf5_chunks.py:636) into ``shm``; returns the
    vchunksx-shaped view."""
    info = tomo_info(filename)
    data = _buffer(shm, tuple(vchunksx))
    z0 = vchunksx[1] * ivchunkx[1]
    z1 = min(z0 + vchunksx[1], info["shape"][1])
    nt, nx = min(vchunksx[0], info["shape"][0]), min(vchunksx[2], info["shape"][2])
    if z1 > z0:
        if (nt, z1 - z0, nx) == tuple(vchunksx):
            read_box(filename, (0, z0, 0), (nt, z1 - z0, nx), out=data)
        else:
            data[:nt, : z1 - z0, :nx] = read_box(filename, (0, z0, 0), (nt, z1 - z0, nx))
    return data


def write_vchunkx(filename, shm, vchunksx, vchunks, ctx, ivchunkx):
    """Writes the vchunksx-shaped buffer ``shm`` as vchunks-sized pieces along z (iohdf5/dxchange_hdf5_chunks.py:681)."""
    data = _buffer(shm, tuple(vchunksx))
    for iblock in range((vchunksx[1] + vchunks[1] - 1) // vchunks[1]):
        b0, b1 = iblock * vchunks[1], min((iblock + 1) * vchunks[1], vchunksx[1])
        if b1 > b0:
            tomo_writex(filename, data[:, b0:b1, :], ivchunk=(ivchunkx[0], iblock, ivchunkx[2]), ctx=ctx)


def write_theta(filename, theta_deg):
    """/exchange/theta (float32 degrees) replaced (step2_radon.py:134-137)."""
    t = np.ascontiguousarray(theta_deg, dtype=np.float32)
    _check(load().mm_h5_write_theta(os.fsencode(filename), t.ctypes.data, t.size), "write_theta")


def read_theta(filename):
    """/exchange/theta as float32 degrees (step3_propagation.py:133)."""
    n = C.c_longlong(0)
    _check(load().mm_h5_read_theta(os.fsencode(filename), None, C.byref(n)), "read_theta")
    t = np.empty(n.value, np.float32)
    _check(load().mm_h5_read_theta(os.fsencode(filename), t.ctypes.data, C.byref(n)), "read_theta")
    return t


def slabs_to_device(filename, axis, slab, device):
    """Yields (start, device tensor) for consecutive slabs of ``slab`` items along ``axis`` (0: projections,
    1: slices) of ``/exchange/data``.  Each slab is read by the HDF5 library straight into a page-locked
    buffer on a reader thread while the previous one uploads and is used (two buffers, the library is
    only ever called from that one thread), so the device never waits on a pageable copy."""
    import torch

    from model_mosaic_amd._carriers import pinned_empty
    info = tomo_info(filename)
    shape = info["shape"]
    bshape = list(shape)
    bshape[axis] = slab
    pins = [pinned_empty(tuple(bshape), torch.float32) for _ in range(2)]
    starts = list(range(0, shape[axis], slab))
    ready = [threading.Event() for _ in starts]
    free = [threading.Event() for _ in range(2)]
    for e in free:
        e.set()
    errors = []

    def reader():
        try:
            for i, s in enumerate(starts):
                free[i % 2].wait()
                free[i % 2].clear()
                cnt = list(shape)
                cnt[axis] = min(slab, shape[axis] - s)
                st = [0, 0, 0]
                st[axis] = s
                view = pins[i % 2].view(-1)[: int(np.prod(cnt))].view(cnt)
                read_box(filename, st, cnt, out=view)
                ready[i].set()
        except Exception as e:  # surfaced to the consumer
            errors.append(e)
            for r in ready:
                r.set()

    th = threading.Thread(target=reader, daemon=True)
    th.start()
    try:
        for i, s in enumerate(starts):
            ready[i].wait()
            if errors:
                raise errors[0]
            cnt = list(shape)
            cnt[axis] = min(slab, shape[axis] - s)
            host = pins[i % 2].view(-1)[: int(np.prod(cnt))].view(cnt)
            dev = host.to(device, non_blocking=False)
            free[i % 2].set()
            yield s, dev
    finally:
        for e in free:
            e.set()
        th.join()
