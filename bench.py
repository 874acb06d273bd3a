"""Benchmark of the USFFT Radon R+RT hot path (and the forward-model iteration) on one MI355X.

Workload (batched FP32 USFFT R and RT on one GPU): a float32 delta volume
nz=64 x 2048 x 2048 of 12 seeded random ellipsoids plus a 1% Gaussian-noise
texture, ntheta=2048 angles uniform on [0, pi), detector 2048, complex64 padded
Fourier grid 4096^2.  One step = R of all 64 slices followed by RT of the
resulting sinogram.  value = slices / s (R+RT) over the timed steps.

Also reported in the same JSON line:
  * roofline of the dominant kernel (HIP events on the launch stream, live over
    the timed region) and of the whole step against the algorithmic bytes
    B_dir(n, ntheta) = 4 n^2 + 32 n (n+1) + 4 ntheta n per slice per direction;
  * per-stage milliseconds;
  * the plain one-slice-at-a-time path (same kernels, nz = 1 per call) and the
    batched/plain ratio;
  * the forward+backward model iteration (nz=128 x 1024^2, ntheta=1024, 30 keV,
    z = 0.2 m, pixel 1.38 um, Poisson data at 1e4 photons/pixel) in s/iter;
  * Fresnel D and DT of 128 complex64 frames 2048 x 2048 (psi = exp(i p - 1e-3 p)
    of a smooth random field p with a 2 rad peak, 30 keV, pixel 0.69 um,
    z in {0.05, 0.2, 0.8} m) in frames/s, against B_D = 4 nz n 8 bytes per frame;
  * cpu_baseline: the NumPy FP32 restatement (oracle) of R+RT on the host
    cores (one thread and all cores of the affinity set, at most 16), for one
    2048^2 slice at 2048 angles and for nz=8 x 256^2 at 384 angles (median
    of 5), with the CPU model.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "usfft_R_RT_n2048_nz64_hbm_traffic.json")


def measured_traffic(stage, n, nt, nz):
    """HBM bytes per launch of a stage, measured offline with two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE) over this workload and scaled by the counts of a plain 1 GiB device copy; None when no
    measurement of this workload is on file."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("workload") != {"n": n, "ntheta": nt, "nz": nz}:
        return None
    st = d.get("stages", {}).get(stage)
    return None if st is None else int(st["traffic_bytes"])


def b_dir(n, nt):
    return 4 * n * n + 32 * n * (n + 1) + 4 * nt * n


def stage_bytes(n, nt):
    """Algorithmic bytes per slice of each stage: its own input read once + output written once."""
    h = n // 2 + 1
    half = 8 * n * (n + 1)       # n rows (or columns) of the half spectrum
    full = 16 * n * (n + 1)      # the 2n x (n+1) half spectrum
    reps = 8 * nt * h            # representative samples
    return {
        "r_rows_rfft": 4 * n * n + half,
        "r_col_fft": half + full,
        "r_gather": full + reps,
        "r_line_ifft": reps + 4 * nt * n,
        "rt_line_fft": 4 * nt * n + reps,
        "rt_scatter": reps + full,
        "rt_col_ifft": full + half,
        "rt_rows_c2r": half + 4 * n * n,
    }


_CPU_JOB = {}


def _cpu_R_block(bounds):
    """R of the whole volume at the angles [b0, b1) (oracle, FP32), rescaled to the full angle count: the angle
    count enters R only through the 1 / sqrt(n ntheta) of phi."""
    import numpy as np

    from oracle import usfft
    b0, b1 = bounds
    x, th = _CPU_JOB["x"], _CPU_JOB["theta"]
    return usfft.tomo_R(x, th[b0:b1], np.float32) * np.float32(math.sqrt((b1 - b0) / len(th)))


def _cpu_RT_block(args):
    """RT of the angles [b0, b1) of a sinogram (oracle, FP32), rescaled like _cpu_R_block; the blocks sum to RT."""
    import numpy as np

    from oracle import usfft
    (b0, b1), sino = args
    th = _CPU_JOB["theta"]
    return usfft.tomo_RT(sino, th[b0:b1], dtype=np.float32) * np.float32(math.sqrt((b1 - b0) / len(th)))


def _cpu_R_RT(pool, workers):
    import numpy as np
    nt = len(_CPU_JOB["theta"])
    edges = np.linspace(0, nt, workers + 1).astype(int)
    blocks = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    if pool is None:
        sino = _cpu_R_block((0, nt))
        return sino, _cpu_RT_block(((0, nt), sino))
    sino = np.concatenate(list(pool.map(_cpu_R_block, blocks)), axis=0)
    back = sum(pool.map(_cpu_RT_block, [(b, sino[b[0]:b[1]]) for b in blocks]))
    return sino, back


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline():
    """The oracle's FP32 NumPy restatement of R+RT on the host cores, run before the GPU is initialised (the
    worker processes are forked).  Two workloads: the small volume nz=8 x 256^2 at 384 angles (one warm-up, median
    of 5 runs, on one thread and on all cores) and one 2048^2 slice at all 2048 angles (one run on all cores; on
    one thread it takes ~40 s).  "All cores" = the cores of this process's affinity set, at most 16: angle blocks
    of R and of RT in worker processes, the RT blocks summed.
    The value is the all-cores rate of the 2048-angle n=2048 slice in slices/s."""
    import concurrent.futures as cf
    import multiprocessing as mp
    import statistics

    import numpy as np
    from threadpoolctl import threadpool_limits

    from model_mosaic_amd import phantom
    from oracle import usfft

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    ctx = mp.get_context("fork")
    res = {"cpu_model": _cpu_model(), "workers": workers}

    def timed(pool, reps):
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            out = _cpu_R_RT(pool, workers)
            times.append(time.perf_counter() - t0)
        return statistics.median(times), out

    with threadpool_limits(limits=1):
        # small volume: 8 slices of the 12-ellipsoid phantom, 384 angles
        _CPU_JOB.update(x=phantom.ellipsoids(8, 256, seed=0).numpy(), theta=phantom.angles(384))
        timed(None, 1)
        t1, (sino, _) = timed(None, 5)
        with cf.ProcessPoolExecutor(workers, mp_context=ctx) as pool:
            timed(pool, 1)
            tw, _ = timed(pool, 5)
        nt_small = len(_CPU_JOB["theta"])
        tr, tt = [], []
        for _ in range(5):  # R and RT apart, one thread
            t0 = time.perf_counter()
            s1 = _cpu_R_block((0, nt_small))
            t_mid = time.perf_counter()
            _cpu_RT_block(((0, nt_small), s1))
            tr.append(t_mid - t0)
            tt.append(time.perf_counter() - t_mid)
        r64 = usfft.tomo_R(_CPU_JOB["x"], _CPU_JOB["theta"], np.float64)
        err = float(np.linalg.norm(sino - r64) / np.linalg.norm(r64))
        res["small"] = {"workload": "nz=8 x 256^2, 384 angles, R+RT", "median_of": 5,
                        "one_thread_s": round(t1, 3), "one_thread_slices_per_s": round(8 / t1, 3),
                        "one_thread_R_slices_per_s": round(8 / statistics.median(tr), 3),
                        "one_thread_RT_slices_per_s": round(8 / statistics.median(tt), 3),
                        "all_cores_s": round(tw, 3), "all_cores_slices_per_s": round(8 / tw, 3),
                        "R_rel_l2_fp32_vs_fp64": err}
        # Fresnel D of one 2048 x 2048 frame (FP32 restatement: pad, FFT2, chirp, IFFT2, crop), one thread
        from oracle import fresnel as fr
        frame = (np.random.default_rng(3).standard_normal((1, 2048, 2048)) * 1j).astype(np.complex64)
        fk = fr.fresnel_kernel_2d(2048, 2048, 1.24e-9 / 30.0, 0.69e-6, [0.2], np.float32)[0]
        fr.D(frame, fk, np.float32)
        tf = []
        for _ in range(3):
            t0 = time.perf_counter()
            fr.D(frame, fk, np.float32)
            tf.append(time.perf_counter() - t0)
        res["fresnel_frame"] = {"workload": "D of one 2048 x 2048 complex64 frame, z = 0.2 m", "median_of": 3,
                                "one_thread_s": round(statistics.median(tf), 3),
                                "one_thread_frames_per_s": round(1.0 / statistics.median(tf), 3)}
        del frame, fk
        # one n=2048 slice at all 2048 angles
        _CPU_JOB.update(x=phantom.ellipsoids(64, 2048, seed=0, noise=0.01, rows=(32, 33)).numpy(), theta=phantom.angles(2048))
        with cf.ProcessPoolExecutor(workers, mp_context=ctx) as pool:
            tw2, _ = timed(pool, 1)
        res["n2048"] = {"workload": "1 slice 2048^2, 2048 angles, R+RT", "all_cores_s": round(tw2, 3)}
    _CPU_JOB.clear()
    return {
        "value": 1.0 / tw2,
        "unit": "slices/s",
        "cores": workers,
        "kind": "port",
        "sample": (f"oracle.usfft FP32 R+RT, one 2048^2 slice at 2048 angles: {tw2:.2f} s on {workers} worker "
                   f"processes (angle blocks); nz=8 x 256^2 at 384 angles: median of 5 "
                   f"{tw:.3f} s on {workers} workers, {t1:.3f} s on one thread; CPU {res['cpu_model']}"),
        "detail": res,
    }


def run_fresnel(torch, reps=3):
    """D and DT of 128 frames of 2048 x 2048 at the three distances; seconds per call, frames/s."""
    from model_mosaic_amd import phantom
    from model_mosaic_amd.processing.propagation import Propagation
    from model_mosaic_amd.step_kernels import psi_from_proj

    b, nz, n = 128, 2048, 2048
    wl, vox, dist = 1.24e-9 / 30.0, 0.69e-6, [0.05, 0.2, 0.8]
    dev = torch.device("cuda", 0)
    p = phantom.smooth_field(b, nz, n, peak=2.0, seed=3, device=dev)
    psi = psi_from_proj(p, 1.0, 1e-3)  # exp(i p - 1e-3 p)
    del p
    out = torch.empty_like(psi)
    prop = Propagation(n, nz, b, len(dist), wl, vox, dist)
    frame_bytes = 4 * nz * n * 8
    res = {"config": f"{b} frames {nz}x{n} c64, 30keV, pixel 0.69um, z={dist} m, symmetric pad",
           "algorithmic_bytes_per_frame": frame_bytes}
    for name, f in (("D", prop.D), ("DT", prop.DT)):
        for j in range(len(dist)):
            f(psi, j, out=out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for j in range(len(dist)):
                f(psi, j, out=out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / (reps * len(dist))
        res[name] = {"ms_per_call": round(dt * 1e3, 4), "frames_per_s": round(b / dt, 1),
                     "achieved_GBs": round(b * frame_bytes / dt / 1e9, 1),
                     "hbm_frac": round(b * frame_bytes / dt / 1e9 / PEAK_HBM_GBS, 4)}
    prop.free()
    return res


def run_paganin(torch, reps=5):
    """Paganin retrieval of 128 intensity frames of 2048 x 2048 (30 keV, 0.69 um, 0.8 m, alpha 1e-3).  Algorithmic
    bytes = the float32 frame read once and written once; the three-pass kernels move 24 bytes per pixel."""
    from model_mosaic_amd.processing.paganin import Paganin
    b, nz, n = 128, 2048, 2048
    dev = torch.device("cuda", 0)
    x = torch.rand((b, nz, n), device=dev, dtype=torch.float32).mul_(0.5).add_(0.5)
    out = torch.empty_like(x)
    pg = Paganin(n, nz, b, 1.24e-9 / 30.0, 0.69e-6, 0.8, 1e-3)
    pg.retrieve(x, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        pg.retrieve(x, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    alg = 8 * b * nz * n
    pg.free()
    del x, out
    torch.cuda.empty_cache()
    return {"config": f"{b} frames {nz}x{n} float32, 30keV, pixel 0.69um, z=0.8 m, alpha=1e-3",
            "ms_per_call": round(dt * 1e3, 4), "frames_per_s": round(b / dt, 1),
            "algorithmic_GBs": round(alg / dt / 1e9, 1), "hbm_frac": round(alg / dt / 1e9 / PEAK_HBM_GBS, 4),
            "moved_GBs": round(3 * alg / dt / 1e9, 1)}


def run_large(torch, reps=2):
    """TomoLargeReal.R + RT of a host (numpy) 64 x 2048 x 2048 volume at 2048 angles, PCIe included: one
    device batch, and 4 z slabs streamed through the three-stream pipeline (budget of 16 slices)."""
    import numpy as np

    from model_mosaic_amd import phantom
    from model_mosaic_amd.processing.tomo_large import TomoLargeReal
    nz, n, nt = 64, 2048, 2048
    th = phantom.angles(nt)
    obj = np.random.default_rng(0).standard_normal((nz, n, n), dtype=np.float32)
    res = {"config": f"host float32 volume {nz}x{n}x{n}, {nt} angles, R then RT, PCIe-inclusive"}
    for name, zb in (("one_batch", nz), ("streamed_16", 16)):
        probe = TomoLargeReal(n, th, n)
        tl = TomoLargeReal(n, th, n, gpu_budget_bytes=zb * probe._slice_bytes(False) + 1)
        sino = tl.R(obj, [n, nt, n]).copy()
        tl.RT(sino, [n, nt, n])
        t0 = time.perf_counter()
        for _ in range(reps):
            sino = tl.R(obj, [n, nt, n])
            tl.RT(sino, [n, nt, n])
        dt = (time.perf_counter() - t0) / reps
        res[name] = {"s_per_RT_R": round(dt, 4), "slices_per_s": round(nz / dt, 1)}
        tl.free()
    torch.cuda.empty_cache()
    return res


def run_model(torch, steps=3, warmup=1):
    import numpy as np

    from model_mosaic_amd import phantom
    from model_mosaic_amd.model import ForwardModel, misfit
    from model_mosaic_amd.step_kernels import abs2, psi_from_proj

    nz, n, nt = 128, 1024, 1024
    wl, vox, z = 1.24e-9 / 30.0, 1.38e-6, 0.2
    theta = phantom.angles(nt)
    dev = torch.device("cuda", 0)
    delta_true = phantom.ellipsoids(nz, n, seed=4, device=dev)
    model = ForwardModel(n, nz, theta, wl, vox, [z])
    with torch.no_grad():
        p = model.tomo.R(delta_true)
        psi = psi_from_proj(p, model.k, 1e-3)          # beta = 1e-3 delta  =>  R beta = 1e-3 R delta
        inten = abs2(model.prop.D(psi, 0))
        g = torch.Generator(device=dev).manual_seed(5)
        data = torch.poisson(inten * 1e4, generator=g) / 1e4
        del p, psi, inten
    delta = (0.9 * delta_true).requires_grad_(True)
    beta = (0.9e-3 * delta_true).requires_grad_(True)

    def it():
        loss = misfit(delta, beta, model, data)
        loss.backward()
        delta.grad = None
        beta.grad = None
        return loss

    for _ in range(warmup):
        it()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = it()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    bdir = b_dir(n, nt)
    px = nt * nz * n
    # yardstick: R of delta and beta and RT of the two gradients (2 directions each, B_dir per slice), a fused
    # psi-build + D + |.|^2 (f32 in, c64 through, f32 out: 24 B/px), D recompute and DT (two separable passes
    # of complex64 each: 32 B/px each) and the chain rule (16 B/px)
    yard = 4 * nz * bdir + px * (24 + 32 + 32 + 16)
    # this implementation: the same four Radon directions; forward = psi-build inside D's column pass (8 B/px
    # in, 8 out) + row pass keeping u and reading data (8 + 8 + 4); backward from the kept p, b, u = residual
    # inside DT's column pass (8 + 4 in, 8 out) + chain rule inside its row pass (8 + 8 in, 8 out)
    impl = 4 * nz * bdir + px * (16 + 20 + 20 + 24)
    model.free()
    return {"s_per_iter": dt, "loss": float(loss.detach()),
            "config": f"nz={nz} n={n} ntheta={nt} 30keV z=0.2m pixel=1.38um",
            "yardstick_GB": yard / 1e9, "hbm_frac": yard / dt / 1e9 / PEAK_HBM_GBS,
            "implementation_GB": impl / 1e9, "implementation_hbm_frac": impl / dt / 1e9 / PEAK_HBM_GBS}


def run_fbp(torch, reps=5):
    """FBP filter (shepp) of a 2048-angle x 64-slice x 2048 sinogram in place: one read + one write.  Timed
    for the default length-2n transforms and for the reference's 4n edge-padded transforms."""
    from model_mosaic_amd.processing.fbp_filter import FBPFilter
    nt, nz, n = 2048, 64, 2048
    dev = torch.device("cuda", 0)
    sino = torch.randn((nt, nz, n), device=dev, dtype=torch.float32)
    nbytes = 2 * 4 * nt * nz * n
    out = {"config": f"shepp filter, sinogram {nt} x {nz} x {n} float32, edge padding to 4n, in place"}
    for method in ("2n", "4n"):
        f = FBPFilter(n, method=method)
        w = f.calc_filter("shepp")
        f.filter(sino, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f.filter(sino, w)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[method] = {"ms_per_call": round(dt * 1e3, 4), "achieved_GBs": round(nbytes / dt / 1e9, 1),
                       "hbm_frac": round(nbytes / dt / 1e9 / PEAK_HBM_GBS, 4)}
    # FBP: filter then RT (two passes over the sinogram) vs RT_filtered (filter inside RT's line pass)
    from model_mosaic_amd import phantom
    from model_mosaic_amd.processing.tomo import TomoReal
    tomo = TomoReal(n, nz, phantom.angles(nt))
    f = FBPFilter(n)
    w = f.calc_filter("shepp")
    rec = torch.empty((nz, n, n), device=dev, dtype=torch.float32)
    work = torch.empty_like(sino)

    def two_pass():
        work.copy_(sino)  # the filter is in place; RT_filtered leaves its input alone
        tomo.RT(f.filter(work, w), out=rec)

    def fused():
        tomo.RT_filtered(sino, f, w, out=rec)

    copy_s = None
    for name, fn in (("copy", lambda: work.copy_(sino)), ("filter_then_RT", two_pass), ("RT_filtered", fused)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        if name == "copy":
            copy_s = dt
            continue
        if name == "filter_then_RT":
            dt -= copy_s  # the copy only keeps the input for the next repetition
        out[name + "_ms"] = round(dt * 1e3, 4)
    tomo.free()
    del sino, work, rec
    torch.cuda.empty_cache()
    return out


def run_mosaic(torch, reps=2):
    """R, psi, Fresnel D (30 keV, 0.69 um, z = 0.2 m) and |.|^2 over the uneven mosaic as one HIP graph."""
    import numpy as np

    from model_mosaic_amd import phantom
    from model_mosaic_amd.mosaic import MosaicBatch

    spec = phantom.uneven_mosaic()
    dev = torch.device("cuda", 0)
    deltas = [phantom.ellipsoids(t["ztotal"], t["n"], seed=100 + t["x"], rows=t["rows"], device=dev) for t in spec]
    tiles = [(t["n"], t["nz"], phantom.angles(t["ntheta"])) for t in spec]
    # one k for the batch: the largest phase k p about 2 rad (p ~ n max(delta) / (2 sqrt(n ntheta)))
    pmax = max(float(d.max()) * 0.5 * np.sqrt(t["n"] / t["ntheta"]) for d, t in zip(deltas, spec))
    k = 2.0 / pmax
    t0 = time.perf_counter()
    mb = MosaicBatch(tiles, 1.24e-9 / 30.0, 0.69e-6, 0.2, k, 1e-3, autotune=True)
    plan_s = time.perf_counter() - t0
    outs = mb.out_buffers()
    mb.run(deltas, outs)                       # autotune + capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        mb.run(deltas, outs)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    info = mb.info()
    # per tile: R (B_dir per slice) + Fresnel D of the ntheta frames at B_D = 4 passes of complex64 (32 B/px)
    alg = sum(t["nz"] * b_dir(t["n"], t["ntheta"]) + 32 * t["ntheta"] * t["nz"] * t["n"] for t in spec)
    nproj = sum(t["ntheta"] for t in spec)
    res = {"config": "4 x-tiles x 9 z-stacks, 96 rows each (8-row overlaps), widths per x-tile from {1536, 2048, 3072, "
                     "4096}, ntheta per tile from {1029, 2058}, R then psi, Fresnel D (30 keV, 0.69 um, 0.2 m), |.|^2",
           "widths": sorted({t["n"] for t in spec}), "tiles": len(spec), "projections": nproj,
           "s_per_batch": round(dt, 5), "projections_per_s": round(nproj / dt, 1),
           "algorithmic_GB": round(alg / 1e9, 3), "hbm_frac": round(alg / dt / 1e9 / PEAK_HBM_GBS, 4),
           "plan_s": round(plan_s, 2), "radon_plans": info["radon_plans"], "fresnel_plans": info["fresnel_plans"],
           "graph_captures": info["graph_captures"], "branches": info["branches"],
           "tuned_per_plan": sorted({(t["n"], t["ntheta"], g, c, r, z) for t, g, c, r, z in
                                     zip(spec, info["gather_xcd_group"], info["col_split"], info["gather_root"],
                                         info["gather_zp"])}),
           "tuned_fields": ["n", "ntheta", "gather_xcd_group", "col_split", "gather_root", "gather_zp"],
           "workspace_GB": round(info["workspace_bytes"] / 1e9, 2)}
    mb.free()
    del deltas, outs
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--ntheta", type=int, default=2048)
    ap.add_argument("--no-model", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fresnel", action="store_true")
    ap.add_argument("--no-mosaic", action="store_true")
    ap.add_argument("--no-fbp", action="store_true")
    ap.add_argument("--no-paganin", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--only-mosaic", action="store_true", help="run only the uneven mosaic leg (testing)")
    ap.add_argument("--gather-zp", type=int, default=2, choices=(1, 2),
                    help="z slices per gather patch element (TomoReal gather_zp)")
    ap.add_argument("--gather-xcd-group", type=int, default=0, help="override the gather XCD group knob")
    ap.add_argument("--scatter-xcd-group", type=int, default=0, help="override the scatter XCD group knob")
    ap.add_argument("--col-split", type=int, default=-1, help="override the column-FFT split knob (0 / 1)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the R+RT plan's launch knobs before the warm-up (at n = 2048 the defaults are "
                         "the measured optimum; the tuner's picks land within run-to-run noise)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run a plain device copy of this many MiB first (PMC byte-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (R and RT shard over z with no "
              "communication between GPUs). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    cpu = None
    if not a.no_cpu and not a.only_mosaic:
        # before the GPU is initialised: the baseline forks worker processes
        print("[bench] cpu baseline (oracle, host cores)", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
        print(f"[bench] cpu baseline {cpu['sample']}", file=sys.stderr, flush=True)

    import numpy as np
    import torch

    from model_mosaic_amd import _lib, phantom
    from model_mosaic_amd.processing.tomo import TomoReal

    lib = _lib.load()
    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    if a.only_mosaic:
        print(json.dumps({"mosaic": run_mosaic(torch)}), flush=True)
        return
    nz, n, nt = a.nz, a.n, a.ntheta
    if a.copy_calib_mib > 0:
        src = torch.ones(a.copy_calib_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {a.copy_calib_mib} MiB x3 done", file=sys.stderr, flush=True)
        del src, dst
    theta = phantom.angles(nt)
    print(f"[bench] building phantom {nz}x{n}x{n}, ntheta={nt}", file=sys.stderr, flush=True)
    obj = phantom.ellipsoids(nz, n, seed=0, noise=0.01, device=dev)
    t0 = time.perf_counter()
    tomo = TomoReal(n, nz, theta, gather_zp=a.gather_zp)
    plan_s = time.perf_counter() - t0
    if a.gather_xcd_group:
        _lib.check(lib.mm_radon_tune(tomo._plan, _lib.MM_RADON_TUNE_GATHER_XCD_GROUP, a.gather_xcd_group), "tune")
    if a.scatter_xcd_group:
        _lib.check(lib.mm_radon_tune(tomo._plan, _lib.MM_RADON_TUNE_SCATTER_XCD_GROUP, a.scatter_xcd_group), "tune")
    if a.col_split >= 0:
        _lib.check(lib.mm_radon_tune(tomo._plan, _lib.MM_RADON_TUNE_COL_SPLIT, a.col_split), "tune")
    sino = torch.empty((nt, nz, n), dtype=torch.float32, device=dev)
    back = torch.empty((nz, n, n), dtype=torch.float32, device=dev)
    tune_s = 0.0
    if a.autotune:
        # launch shapes only (the results are the same bits): untimed, before the warm-up
        t0 = time.perf_counter()
        _lib.check(lib.mm_radon_autotune(tomo._plan, obj.data_ptr(), sino.data_ptr(), back.data_ptr(), nz,
                                         _lib.stream_handle()), "autotune")
        torch.cuda.synchronize()
        tune_s = time.perf_counter() - t0
    knob_names = {"gather_xcd_group": _lib.MM_RADON_TUNE_GATHER_XCD_GROUP,
                  "scatter_xcd_group": _lib.MM_RADON_TUNE_SCATTER_XCD_GROUP,
                  "col_split": _lib.MM_RADON_TUNE_COL_SPLIT, "gather_root": _lib.MM_RADON_TUNE_GATHER_ROOT,
                  "gather_zp": _lib.MM_RADON_TUNE_GATHER_ZP}
    knobs = {k: int(lib.mm_radon_get_tune(tomo._plan, v)) for k, v in knob_names.items()}
    print(f"[bench] launch knobs {knobs}" + (f" (autotuned in {tune_s:.1f} s)" if a.autotune else ""),
          file=sys.stderr, flush=True)
    print(f"[bench] plan {plan_s:.2f} s, workspace {tomo.workspace_bytes() / 2**30:.2f} GiB; warmup",
          file=sys.stderr, flush=True)

    def step():
        tomo.R(obj, out=sino)
        tomo.RT(sino, out=back)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: per-stage HIP events recorded on the launch stream for every step
    _lib.check(lib.mm_radon_profile(tomo._plan, 1), "profile")
    nst = _lib.MM_RADON_NSTAGES
    acc = np.zeros(nst)
    buf = (np.ctypeslib.ctypes.c_float * nst)()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
        _lib.check(lib.mm_radon_stage_ms(tomo._plan, buf, nst), "stage_ms")
        acc += np.array(buf[:])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    _lib.check(lib.mm_radon_profile(tomo._plan, 0), "profile")
    ms_step = elapsed / a.steps * 1e3
    value = nz * a.steps / elapsed
    stage_ms = {name: acc[i] / a.steps for i, name in enumerate(_lib.RADON_STAGES)}
    sb = stage_bytes(n, nt)
    dom = max(stage_ms, key=stage_ms.get)
    dom_bytes = sb[dom] * nz
    dom_ach = dom_bytes / (stage_ms[dom] * 1e-3) / 1e9
    step_bytes = 2 * nz * b_dir(n, nt)
    step_ach = step_bytes / (ms_step * 1e-3) / 1e9
    print(f"[bench] {value:.1f} slices/s, {ms_step:.2f} ms/step; stages " +
          ", ".join(f"{k}={v:.3f}" for k, v in stage_ms.items()), file=sys.stderr, flush=True)

    plain = plain_ref = None
    if not a.no_plain:
        t1 = TomoReal(n, 1, theta)
        s1 = torch.empty((nt, 1, n), dtype=torch.float32, device=dev)
        o1 = torch.empty((1, n, n), dtype=torch.float32, device=dev)
        k = min(8, nz)
        for z in range(2):
            t1.R(obj[z:z + 1], out=s1)
            t1.RT(s1, out=o1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for z in range(k):
            t1.R(obj[z:z + 1], out=s1)
            t1.RT(s1, out=o1)
        torch.cuda.synchronize()
        plain = k / (time.perf_counter() - t0)
        t1.free()
        print(f"[bench] same kernels one slice per call {plain:.1f} slices/s", file=sys.stderr, flush=True)
        # reference-structured one-slice path: rocFFT via torch.fft + per-sample global-memory gather / atomic scatter
        from model_mosaic_amd.plain_path import PlainTomo
        pt = PlainTomo(n, theta)
        kr = 3
        for z in range(1):
            pt.RT(pt.R(obj[z]))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for z in range(kr):
            pt.RT(pt.R(obj[z]))
        torch.cuda.synchronize()
        plain_ref = kr / (time.perf_counter() - t0)
        del pt
        torch.cuda.empty_cache()
        print(f"[bench] reference-structured one-slice path {plain_ref:.2f} slices/s", file=sys.stderr, flush=True)
    tomo.free()
    del obj, sino, back
    torch.cuda.empty_cache()

    model = None
    if not a.no_model:
        print("[bench] forward+backward model iteration", file=sys.stderr, flush=True)
        model = run_model(torch)
        print(f"[bench] model {model['s_per_iter'] * 1e3:.1f} ms/iter", file=sys.stderr, flush=True)

    fres = None
    if not a.no_fresnel:
        print("[bench] Fresnel D/DT, 128 frames 2048^2", file=sys.stderr, flush=True)
        fres = run_fresnel(torch)
        print(f"[bench] D {fres['D']['ms_per_call']:.2f} ms, DT {fres['DT']['ms_per_call']:.2f} ms per 128 frames",
              file=sys.stderr, flush=True)
        torch.cuda.empty_cache()

    large = None
    if not a.no_large:
        print("[bench] TomoLargeReal R+RT from host arrays", file=sys.stderr, flush=True)
        large = run_large(torch)
        print(f"[bench] host R+RT {large['one_batch']['slices_per_s']:.0f} slices/s one batch, "
              f"{large['streamed_16']['slices_per_s']:.0f} slices/s in 16-slice slabs", file=sys.stderr, flush=True)

    pag = None
    if not a.no_paganin:
        print("[bench] Paganin retrieval, 128 frames 2048^2", file=sys.stderr, flush=True)
        pag = run_paganin(torch)
        print(f"[bench] Paganin {pag['ms_per_call']:.2f} ms per 128 frames ({pag['moved_GBs']:.0f} GB/s moved)",
              file=sys.stderr, flush=True)

    fbp = None
    if not a.no_fbp:
        print("[bench] FBP filter", file=sys.stderr, flush=True)
        fbp = run_fbp(torch)
        print(f"[bench] FBP filter {fbp['2n']['ms_per_call']:.2f} ms (2n transforms), {fbp['4n']['ms_per_call']:.2f} ms "
              f"(4n) per 2048x64x2048 sinogram; filter + RT {fbp['filter_then_RT_ms']:.2f} ms, RT_filtered "
              f"{fbp['RT_filtered_ms']:.2f} ms", file=sys.stderr, flush=True)

    mosaic = None
    if not a.no_mosaic:
        print("[bench] uneven mosaic batch, one HIP graph", file=sys.stderr, flush=True)
        mosaic = run_mosaic(torch)
        print(f"[bench] mosaic {mosaic['s_per_batch'] * 1e3:.1f} ms per batch of {mosaic['tiles']} tiles, "
              f"{mosaic['projections_per_s']:.0f} projections/s", file=sys.stderr, flush=True)

    out = {
        "metric": "USFFT R+RT slices/s @n=2048 & fwd+bwd model s/iter; % HBM roofline",
        "value": round(value, 3),
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched FP32 USFFT R+RT, seeded ellipsoids + 1% noise",
                   "nz": nz, "n": n, "ntheta": nt, "oversampling": 2, "m": 4, "eps": 1e-3,
                   "launch_knobs": knobs, "autotuned": bool(a.autotune)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_ach, 2), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(dom_ach / PEAK_HBM_GBS, 4),
                     "traffic": measured_traffic(dom, n, nt, nz),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per launch, copy-calibrated "
                                       "(profiles/usfft_R_RT_n2048_nz64_hbm_traffic.md)",
                     "algorithmic_bytes": dom_bytes, "avg_ms": round(stage_ms[dom], 4)},
        "step_roofline": {"algorithmic_bytes": step_bytes, "achieved": round(step_ach, 2), "unit": "GB/s",
                          "frac": round(step_ach / PEAK_HBM_GBS, 4)},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "plain_one_slice_slices_per_s": None if plain_ref is None else round(plain_ref, 3),
        "batched_over_plain": None if plain_ref is None else round(value / plain_ref, 2),
        "plain_path": "reference-structured, one slice per call: rocFFT (torch.fft) + per-sample global-memory "
                      "81-tap gather and float-atomic scatter into the full 2n x 2n grid (plain_path.PlainTomo)",
        "same_kernels_one_slice_slices_per_s": None if plain is None else round(plain, 3),
        "batched_over_same_kernels_one_slice": None if plain is None else round(value / plain, 3),
        "model_fwd_bwd": model,
        "fresnel_D_DT": fres,
        "paganin": pag,
        "large_host_R_RT": large,
        "mosaic": mosaic,
        "fbp_filter": fbp,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
