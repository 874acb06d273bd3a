"""bench.py's contract on a machine without a GPU: one JSON line per run is produced only on the GPU, but the
refusal of a multi-GPU request, the algorithmic byte model and the lookup of the measured traffic run here."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and r.stdout == ""
    assert "one GPU" in r.stderr


def test_step_bytes_are_two_directions_of_b_dir():
    """B_dir(n, ntheta) = 4 n^2 (object) + 32 n (n + 1) (the 2n x (n + 1) half spectrum, written and read
    once) + 4 ntheta n (sinogram); the step's yardstick is R and RT of nz slices."""
    n, nt = 2048, 2048
    assert bench.b_dir(n, nt) == 4 * n * n + 32 * n * (n + 1) + 4 * nt * n
    assert 2 * 64 * bench.b_dir(n, nt) == 21483225088


def test_stage_bytes_read_and_write_each_buffer_once():
    """Every stage reads its input and writes its output once; R and RT mirror each other stage by stage."""
    s = bench.stage_bytes(2048, 2048)
    for a, b in (("r_rows_rfft", "rt_rows_c2r"), ("r_col_fft", "rt_col_ifft"), ("r_gather", "rt_scatter"),
                 ("r_line_ifft", "rt_line_fft")):
        assert s[a] == s[b], (a, b)
    assert s["rt_scatter"] * 64 == 5371854848


def test_measured_traffic_only_for_its_own_workload():
    with open(bench.TRAFFIC_FILE) as f:
        d = json.load(f)
    w = d["workload"]
    got = bench.measured_traffic("rt_scatter", w["n"], w["ntheta"], w["nz"])
    assert got == int(d["stages"]["rt_scatter"]["traffic_bytes"]) and got > 0
    assert bench.measured_traffic("rt_scatter", w["n"] // 2, w["ntheta"], w["nz"]) is None
    assert bench.measured_traffic("no_such_stage", w["n"], w["ntheta"], w["nz"]) is None
