"""Builds libmosaic_hip.so (gfx950) in-tree from the HIP sources in csrc/.

Usage: python -m model_mosaic_amd.build [--jobs N]

Every translation unit is compiled with hipcc for gfx950 only, packed-FP32
instruction selection disabled, then linked into one shared library that the
ctypes layer (model_mosaic_amd/_lib.py) loads.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJDIR = HERE / "_build"
LIB = HERE / "libmosaic_hip.so"
SOURCES = ["common.hip", "radon.hip", "fresnel.hip", "model.hip", "mosaic.hip", "plain.hip", "fbp.hip",
           "paganin.hip"]
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 VALU ops are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str) -> Path:
    obj = OBJDIR / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


H5_SRC = CSRC / "h5io.cpp"
H5_LIB = HERE / "libmosaic_h5.so"
H5_PREFIX = Path(os.environ.get("MM_HDF5_PREFIX", "/opt/conda"))


def build_h5(verbose: bool = False):
    """Builds libmosaic_h5.so (the dxchange HDF5 layer, g++ on the HDF5 C library); None if the library's
    headers are not installed (the HDF5 layer then raises when used)."""
    if not (H5_PREFIX / "include" / "hdf5.h").exists():
        if verbose:
            print(f"HDF5 headers not found under {H5_PREFIX}: libmosaic_h5.so not built")
        return None
    tmp = H5_LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", str(H5_SRC),
           f"-I{H5_PREFIX / 'include'}", f"-L{H5_PREFIX / 'lib'}", "-lhdf5", f"-Wl,-rpath,{H5_PREFIX / 'lib'}",
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for h5io.cpp:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print(f"built {H5_LIB}")
    return H5_LIB


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compiles every HIP source and links libmosaic_hip.so, then the HDF5 layer; returns the HIP library's path."""
    OBJDIR.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
