"""The dxchange HDF5 layer (model_mosaic_amd/iohdf5, libmosaic_h5.so) on the host: the file layout of
iohdf5/dxchange_hdf5_chunks.py (virtual master over bank files, writer attributes, theta in degrees),
checked through this layer's own reads and, independently, through the HDF5 command-line tools of the
image (h5dump prints what the stock library reads from the files; h5import writes a file this layer
must read).  No GPU."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

H5BIN = "/opt/conda/bin"


@pytest.fixture(scope="module")
def dx():
    from model_mosaic_amd import build
    from model_mosaic_amd.iohdf5 import dxchange_hdf5_chunks as mod
    if not mod.LIB_PATH.exists() and build.build_h5() is None:
        pytest.skip("HDF5 C library not installed")
    return mod


def h5dump_values(path, tmp, dset="/exchange/data"):
    """A dataset's values as the stock h5dump writes them out (little-endian binary float32)."""
    out = os.path.join(str(tmp), "dump.bin")
    subprocess.run([os.path.join(H5BIN, "h5dump"), "-d", dset, "-b", "LE", "-o", out, path], capture_output=True,
                   check=True)
    return np.fromfile(out, dtype="<f4")


needs_tools = pytest.mark.skipif(not os.path.exists(os.path.join(H5BIN, "h5dump")), reason="HDF5 tools not installed")


def test_proj_banks_round_trip_and_layout(dx, tmp_path):
    """stype 'proj', vchunks of 4 projections in 2 banks over 10 projections (a ragged last vchunk, an empty
    last bank): written vchunk by vchunk, read back as arbitrary boxes and projection slabs, bit for bit."""
    f = str(tmp_path / "proj.h5")
    shape, vchunks = (10, 6, 5), (4, 6, 5)
    ctx = dx.tomo_initx(f, shape, np.float32, vchunks=vchunks, stype="proj", nbanks=2)
    assert ctx["banks_size"] == [2, 2, 2, 2, 2, 0]
    on_disk = sorted(os.listdir(tmp_path / "proj"))
    assert on_disk == [f"proj_data_{i:06d}.h5" for i in range(5)]          # no file for the empty bank
    data = np.random.default_rng(0).standard_normal(shape).astype(np.float32)
    for iv in range(3):
        dx.tomo_writex(f, data[4 * iv:4 * iv + 4], ivchunk=(iv, 0, 0), ctx=ctx)
    info = dx.tomo_info(f)
    assert info["shape"] == shape and info["is_virtual"] and info["stype"] == "proj"
    assert info["vchunks"] == vchunks and info["nbanks_per_svchunk"] == 2 and info["chunks"] == (1, 6, 5)
    assert np.array_equal(dx.read_box(f, (0, 0, 0), shape), data)
    assert np.array_equal(dx.read_box(f, (3, 2, 1), (5, 3, 4)), data[3:8, 2:5, 1:5])
    buf = np.zeros(4 * 6 * 5, np.float32)
    slab = dx.read_projs_vchunkx(f, buf, 2, (4, 6, 5), (2, 0, 0))           # the 2 projections of the last vchunk
    assert np.array_equal(slab[:2], data[8:10])
    with pytest.raises(ValueError):
        dx.read_box(f, (8, 0, 0), (3, 6, 5))                              # outside the dataset


def test_slice_banks_vchunkx_and_theta(dx, tmp_path):
    """stype 'slice' (banks along z): a z-slab buffer written with write_vchunkx in vchunk pieces and read
    back with read_slices_vchunkx; default chunks (vchunks[0], 1, vchunks[2]); theta in degrees."""
    f = str(tmp_path / "data.h5")
    shape, vchunks = (7, 12, 8), (7, 4, 8)
    ctx = dx.tomo_initx(f, shape, np.float32, vchunks=vchunks, stype="slice", nbanks=2)
    data = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    # a buffer over all of z, written as vchunk pieces 0, 1, 2 along z (the pieces of
    # iohdf5/dxchange_hdf5_chunks.py:681-698 are numbered from z = 0 within the buffer)
    dx.write_vchunkx(f, np.ascontiguousarray(data), (7, 12, 8), vchunks, ctx, (0, 0, 0))
    info = dx.tomo_info(f)
    assert info["stype"] == "slice" and info["chunks"] == (7, 1, 8)
    got = dx.read_slices_vchunkx(f, np.zeros(7 * 8 * 8, np.float32), 1, (7, 8, 8), (0, 1, 0))
    assert np.array_equal(got[:, :4], data[:, 8:12])
    assert np.array_equal(dx.read_box(f, (0, 0, 0), shape), data)
    theta = np.rad2deg(np.linspace(0, np.pi, 7, endpoint=False)).astype(np.float32)
    dx.write_theta(f, theta)
    dx.write_theta(f, theta)                                                # replaced, like step2_radon.py:135-137
    assert np.array_equal(dx.read_theta(f), theta)


@needs_tools
def test_files_read_by_the_stock_hdf5_tools(dx, tmp_path):
    """h5dump (the HDF5 library's own reader) sees a virtual /exchange/data over the bank files with the
    writer's attributes and the values written; the master set moved to another directory still resolves."""
    f = str(tmp_path / "proj.h5")
    shape = (5, 3, 4)
    ctx = dx.tomo_initx(f, shape, np.float32, vchunks=(2, 3, 4), stype="proj", nbanks=1)
    data = (np.arange(np.prod(shape)) * 0.25 - 3).astype(np.float32).reshape(shape)
    for iv in range(3):
        dx.tomo_writex(f, data[2 * iv:2 * iv + 2], ivchunk=(iv, 0, 0), ctx=ctx)
    hdr = subprocess.run([os.path.join(H5BIN, "h5dump"), "-A", "-p", f], capture_output=True, text=True,
                         check=True).stdout
    assert "VIRTUAL" in hdr and 'FILE "proj/proj_data_000002.h5"' in hdr
    for attr, val in (("nbanks_per_svchunk", "1"), ("vchunks_0", "2"), ("chunks_0", "1"), ('stype', '"proj"')):
        m = re.search(r'ATTRIBUTE "%s" \{.*?DATA \{\s*\(0\): ([^\s]+)' % attr, hdr, re.S)
        assert m and m.group(1) == val, attr
    assert np.array_equal(h5dump_values(f, tmp_path), data.ravel())
    moved = tmp_path / "moved"
    moved.mkdir()
    shutil.move(f, moved / "proj.h5")
    shutil.move(str(tmp_path / "proj"), moved / "proj")
    assert np.array_equal(dx.read_box(str(moved / "proj.h5"), (0, 0, 0), shape), data)


@needs_tools
def test_reads_a_file_written_by_h5import(dx, tmp_path):
    """A plain (contiguous) /exchange/data written by the stock h5import is read back exactly; no writer
    attributes: one bank, vchunks = shape, no chunks."""
    shape = (5, 4, 3)
    data = np.random.default_rng(2).standard_normal(shape).astype(np.float32)
    raw = tmp_path / "in.bin"
    data.tofile(raw)
    conf = tmp_path / "conf.txt"
    conf.write_text("PATH /exchange/data\nINPUT-CLASS FP\nINPUT-SIZE 32\nRANK 3\nDIMENSION-SIZES 5 4 3\n"
                    "OUTPUT-CLASS FP\nOUTPUT-SIZE 32\nOUTPUT-ARCHITECTURE IEEE\nOUTPUT-BYTE-ORDER LE\n")
    out = str(tmp_path / "imp.h5")
    subprocess.run([os.path.join(H5BIN, "h5import"), str(raw), "-c", str(conf), "-o", out], check=True)
    info = dx.tomo_info(out)
    assert info["shape"] == shape and not info["is_virtual"] and info["chunks"] is None
    assert info["nbanks_per_svchunk"] == 1 and info["vchunks"] == shape
    assert np.array_equal(dx.read_box(out, (0, 0, 0), shape), data)
    dx.write_box(out, (1, 1, 1), np.full((2, 2, 2), 7, np.float32))       # a plain dataset is written in place
    data[1:3, 1:3, 1:3] = 7
    assert np.array_equal(dx.read_box(out, (0, 0, 0), shape), data)


def test_header_symbols_exported(dx):
    """Every entry point include/mosaic_h5.h declares is exported by libmosaic_h5.so and bound here."""
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "mosaic_h5.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(mm_h5_[a-z_]+)\s*\(", hdr))
    lib = dx.load()
    assert names and all(hasattr(lib, n) for n in names)
