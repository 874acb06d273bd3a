"""The HDF5 layer (csrc/h5io.cpp) built under AddressSanitizer (host code, leak detection on): a driver creates
projection- and slice-banked virtual datasets, writes, reads a box back, reads the layout and theta, and takes
two error paths; no sanitizer report and the values read back equal the ones written."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
H5 = Path(os.environ.get("MM_HDF5_PREFIX", "/opt/conda"))


@pytest.mark.skipif(shutil.which("g++") is None or not (H5 / "include" / "hdf5.h").exists(),
                    reason="needs g++ and the HDF5 C library headers")
def test_hdf5_layer_under_address_sanitizer(tmp_path):
    exe = tmp_path / "h5_asan"
    cmd = ["g++", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-std=c++17",
           str(ROOT / "tests" / "h5_asan_driver.cpp"), str(ROOT / "model_mosaic_amd" / "csrc" / "h5io.cpp"),
           f"-I{ROOT / 'include'}", f"-I{H5 / 'include'}", f"-L{H5 / 'lib'}", "-lhdf5", f"-Wl,-rpath,{H5 / 'lib'}",
           "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0")
    run = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, env=env, timeout=120)
    assert run.returncode == 0 and "AddressSanitizer" not in run.stderr, run.stderr[-2000:]
    lines = run.stdout.strip().splitlines()
    assert len(lines) == 2
    for line in lines:
        assert "mismatches 0, virtual 1, shape 20 12 16, theta n 20 equal 1, errors -1 -2" in line, line
