"""Benchmark of the w-stacking imaging hot path on one MI355X.

A step is one pass of the hot path over one batch of synthetic samples already
resident in HBM: bucket + grid -> per-plane FFT -> w-phase correction + plane sum
(lofar_image_dev).  The headline workload is the FP64 drop-in-mode configuration:
16M samples, 4096x4096 grid, 32 w-planes, support 7.  The same run reports the
FP32C precision mode on 64M samples (same grid), the adjoint (degridding) rate,
image errors, the dominant kernel's roofline fraction (HIP events on the stream
the kernels run on) and the reference CPU path timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("Mvis/s end-to-end imaging, FP32-comp vs FP64 at stated image RMS err; adjoint Mvis/s")


def per_kernel_bytes(n, nx, ny, nw, cell_bytes):
    """Algorithmic HBM bytes per launch: B = 36 N + 4 C + 16 Nx Ny split by kernel."""
    C = nw * nx * ny * cell_bytes
    return {"grid_tiles": 36 * n + C, "fft_cols": 2 * C, "rows_phase": C + 16 * nx * ny,
            "total": 36 * n + 4 * C + 16 * nx * ny}


def time_steps(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def profile_stages(fn, reps):
    import torch
    from lofar_amd import imager
    imager.profile_read()
    imager.profile_enable(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    imager.profile_enable(False)
    return {k: (ms / c, c // reps) for k, (ms, c) in imager.profile_read().items()}


def rel_err(a_re, a_im, b_re, b_im):
    import torch
    num = torch.sqrt(torch.sum((a_re - b_re) ** 2 + (a_im - b_im) ** 2))
    den = torch.sqrt(torch.sum(b_re ** 2 + b_im ** 2))
    mx = torch.max(torch.sqrt((a_re - b_re) ** 2 + (a_im - b_im) ** 2))
    mref = torch.max(torch.sqrt(b_re ** 2 + b_im ** 2))
    return float(num / den), float(mx / mref)


def bench_dropin(d, n, nx, ny, nw, support):
    """The reference's own call chain on host arrays (gridding.c -> wstack(),
    fourier_transform.c -> fftw_data()'s plane transform, phase_correction()), FP64, on the
    headline workload: PCIe-inclusive, one warm run then one timed run."""
    import numpy as np
    from lofar_amd import imager
    h = {k: d[k].cpu().numpy() for k in ("uu", "vv", "ww", "vis_real", "vis_img", "weight")}
    t = {}
    for rep in range(2):
        grid = np.zeros(2 * nw * ny * nx)
        gridss = np.empty_like(grid)
        re, im = np.zeros(nx * ny), np.zeros(nx * ny)
        t0 = time.perf_counter()
        imager.wstack(nw, n, 1, 1, h["uu"], h["vv"], h["ww"], h["vis_real"], h["vis_img"],
                      h["weight"], 1.0 / nx, 1.0 / nw, support, nx, ny, grid)
        t1 = time.perf_counter()
        imager.fftw_data(grid, gridss, nx, ny, nw)
        t2 = time.perf_counter()
        imager.phase_correction(gridss, re, im, nx, ny, nw, nx, ny, d["resolution"], d["wmin"],
                                d["wmax"])
        t3 = time.perf_counter()
        t = {"wstack_ms": (t1 - t0) * 1e3, "fftw_data_ms": (t2 - t1) * 1e3,
             "phase_correction_ms": (t3 - t2) * 1e3}
        del grid, gridss
    total = sum(t.values())
    print("bench: drop-in chain %.1f ms (host arrays, PCIe included)" % total, file=sys.stderr,
          flush=True)
    return dict(workload="wstack -> fftw_data -> phase_correction on host arrays, FP64, 16M "
                         "samples, 4096x4096, 32 w-planes (8.6 GB host cube each way)",
                value=n / (total * 1e-3) / 1e6, unit="Mvis/s", ms=total,
                **{k: round(v, 1) for k, v in t.items()})


FIELDS = ("uu", "vv", "ww", "vis_real", "vis_img", "weight")


def save_workload(work, name, d, nx, nw, support, img_re, img_im):
    """Samples and the GPU image of one workload, for the cpu_baseline leg."""
    import numpy as np
    np.savez(os.path.join(work, name + ".npz"), **{k: d[k].cpu().numpy() for k in FIELDS},
             image=img_re.cpu().numpy() + 1j * img_im.cpu().numpy(), nx=nx, ny=nx, nw=nw,
             support=support, resolution=d["resolution"], wmin=d["wmin"], wmax=d["wmax"])


def cpu_baseline(work, names):
    """The reference CPU path on this host (oracle/cpu_baseline.py, a child process): its
    timing on the first workload, and every workload's GPU image error against it."""
    env = dict(os.environ)
    env["MKL_NUM_THREADS"] = "1"  # the FFT leg is a single-threaded FFTW_ESTIMATE plan
    try:
        out = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"),
                              "--work", work, "--names", ",".join(names)],
                             capture_output=True, text=True, env=env, timeout=900)
        sys.stderr.write(out.stderr[-2000:])
        return json.loads(out.stdout.strip().splitlines()[-1])
    except Exception as e:  # reported, never fatal
        return {"error": repr(e)}


# 64-bit LDS atomics per clock per CU that the LDS's 128 bytes/clock/CU could carry; MI355X:
# 256 CUs at 2.4 GHz.  A bandwidth-derived ceiling: AMD publishes no LDS atomic rate.
LDS_ATOMIC_PEAK_G = 16 * 256 * 2.4


def lds_atomics(n, support, prof):
    """Secondary bound of the gridder: 64-bit LDS atomic adds it needs (re and im of every
    cell a sample's window covers, 2*S^2 per sample) per second of grid_tiles time."""
    if "grid_tiles" not in prof:
        return None
    ms = prof["grid_tiles"][0] / max(prof["grid_tiles"][1], 1)
    ops = 2 * support * support * n
    g = ops / (ms * 1e-3) / 1e9
    return {"kernel": "grid_tiles", "algorithmic_atomics": ops, "avg_launch_ms": ms,
            "achieved": g, "peak": LDS_ATOMIC_PEAK_G, "unit": "G 64-bit LDS atomics/s",
            "frac": g / LDS_ATOMIC_PEAK_G,
            "peak_basis": "16 per clock per CU (128 B/clock/CU LDS), 256 CUs, 2.4 GHz"}


TRAFFIC_SOURCE = ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                  "this workload (scaled by the copy calibration), committed beside the kernel "
                  "profiles; not measured in this run")


def bench_fp32c(dev, a, fp64_value, work, names, nx=4096, nw=32, support=7):
    """FP32C imaging of 64M samples on the headline grid; saves the samples and the image
    for the cpu_baseline leg when `work` is given."""
    import torch
    from lofar_amd import imager
    from lofar_amd.synth import make_obs
    ny = nx
    n3 = 64 * 1024 * 1024
    d3 = make_obs(n3, nx, nw, seed=2025, backend="torch", device=dev)
    vis3 = [d3[k] for k in FIELDS]
    cube32 = torch.empty((nw, ny, nx), dtype=torch.complex64, device=dev)
    re32 = torch.empty((ny, nx), dtype=torch.float64, device=dev)
    im32 = torch.empty_like(re32)

    def step32():
        imager.image(*vis3, nw, support, nx, ny, d3["resolution"], d3["wmin"], d3["wmax"],
                     precision="fp32c", cube=cube32, image_real=re32, image_imag=im32)

    ms32 = time_steps(step32, a.steps, a.warmup)
    prof32 = profile_stages(step32, 3)
    mvis32 = n3 / (ms32 * 1e-3) / 1e6
    kb = per_kernel_bytes(n3, nx, ny, nw, 8)
    out = {"workload": "64M samples, 4096x4096, 32 w-planes, support 7, complex64 cube",
           "value": mvis32, "unit": "Mvis/s", "ms_per_step": ms32,
           "speedup_vs_fp64_mode": (mvis32 / fp64_value) if fp64_value else None,
           "end_to_end_algorithmic_GBs": kb["total"] / (ms32 * 1e-3) / 1e9,
           "stages_ms": {k: round(v[0], 4) for k, v in prof32.items()},
           "gridder_lds_atomics": lds_atomics(n3, support, prof32)}
    print("bench: fp32c %.3f ms/step %.1f Mvis/s" % (ms32, mvis32), file=sys.stderr, flush=True)
    if work is not None:
        save_workload(work, "fp32c", d3, nx, nw, support, re32, im32)
        names.append("fp32c")
    del d3, vis3, cube32, re32, im32
    torch.cuda.empty_cache()
    return out


def bench_adjoint(d, vis, cubeA, n2, nx, ny, nw, support, a):
    """Degridding a 4096^2 image to the 16M samples (FP32C), and the dot test
    <y, A x> = <A^H y, x> on that workload."""
    import torch
    from lofar_amd import imager
    dev = cubeA.device
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    img_re = torch.randn((ny, nx), generator=g, dtype=torch.float64, device=dev)
    img_im = torch.zeros_like(img_re)
    mr = torch.empty(n2, dtype=torch.float32, device=dev)
    mi = torch.empty_like(mr)

    def stepA():
        imager.degrid(img_re, img_im, d["uu"], d["vv"], d["ww"], d["weight"], nw, support,
                      d["resolution"], d["wmin"], d["wmax"], precision="fp32c", cube=cubeA,
                      model_real=mr, model_imag=mi)

    msA = time_steps(stepA, a.steps, a.warmup)
    profA = profile_stages(stepA, 3)
    # A x = the FP32C dirty image of the 16M visibilities, A^H y = the FP32C degridding of
    # y = A x (normalised) + noise
    ar, ai = imager.image(*vis, nw, support, nx, ny, d["resolution"], d["wmin"], d["wmax"],
                          precision="fp32c", cube=cubeA)
    rms = torch.sqrt(torch.mean(ar * ar + ai * ai))
    yr = ar / rms + torch.randn((ny, nx), generator=g, dtype=torch.float64, device=dev)
    yi = ai / rms + torch.randn((ny, nx), generator=g, dtype=torch.float64, device=dev)
    imager.degrid(yr, yi, d["uu"], d["vv"], d["ww"], d["weight"], nw, support,
                  d["resolution"], d["wmin"], d["wmax"], precision="fp32c", cube=cubeA,
                  model_real=mr, model_imag=mi)
    xr, xi = d["vis_real"].double(), d["vis_img"].double()
    lhs_re = float((yr * ar + yi * ai).sum())
    lhs_im = float((yr * ai - yi * ar).sum())
    hr, hi = mr.double(), mi.double()
    rhs_re = float((hr * xr + hi * xi).sum())
    rhs_im = float((hr * xi - hi * xr).sum())
    dot_rel = (math.hypot(lhs_re - rhs_re, lhs_im - rhs_im) / math.hypot(lhs_re, lhs_im))
    print("bench: adjoint %.3f ms/step, dot test %.2e" % (msA, dot_rel), file=sys.stderr,
          flush=True)
    C = nw * nx * ny * 8
    alg = 16 * nx * ny + 4 * C + 44 * n2
    return {"workload": "degrid 4096x4096 image to 16M samples, 32 w-planes, FP32C",
            "value": n2 / (msA * 1e-3) / 1e6, "unit": "Mvis/s", "ms_per_step": msA,
            "algorithmic_GBs": alg / (msA * 1e-3) / 1e9, "dot_test_rel_err": dot_rel,
            "stages_ms": {k: round(v[0], 4) for k, v in profA.items()}}


def bench_chi2(d, n2, nx, ny, nw, support):
    """chi^2 sky-model fit: 20 gradient-descent iterations through the autograd binding,
    each one degrid (forward) + one imaging of the weighted residuals (backward)."""
    import torch
    from lofar_amd.autograd import ImagingOperator, chi2_descent, lipschitz
    dev = d["uu"].device
    g = torch.Generator(device=dev)
    g.manual_seed(6)
    op = ImagingOperator(d["uu"], d["vv"], d["ww"], None, nw, support, nx, ny,
                         d["resolution"], d["wmin"], d["wmax"], precision="fp32c")
    yy, xx = torch.meshgrid(torch.arange(ny, device=dev, dtype=torch.float64),
                            torch.arange(nx, device=dev, dtype=torch.float64), indexing="ij")
    sky = 1e-3 * torch.randn((ny, nx), generator=g, dtype=torch.float64, device=dev)
    for (cy, cx, amp, sig) in ((1200.0, 900.0, 1.0, 3.0), (2900.0, 2500.0, 0.6, 5.0),
                               (700.0, 3300.0, 1.5, 2.0)):
        sky += amp * torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))
    del yy, xx
    zero = torch.zeros_like(sky)
    step = 0.5 / lipschitz(op, d["weight"], iters=4)
    hist = []
    chi2_descent(op, d["vis_real"], d["vis_img"], d["weight"], sky, zero, iters=1, step=step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chi2_descent(op, d["vis_real"], d["vis_img"], d["weight"], sky, zero, iters=20, step=step,
                 history=hist)
    torch.cuda.synchronize()
    msG = (time.perf_counter() - t0) * 1e3 / 20
    h = [float(v) for v in hist]
    print("bench: chi2 fit %.3f ms/iteration, chi2 %.4e -> %.4e" % (msG, h[0], h[-1]),
          file=sys.stderr, flush=True)
    return {"workload": "20 gradient-descent iterations of the weighted chi^2 of a 4096x4096 "
                        "sky model against 16M samples, 32 w-planes, FP32C, torch.autograd "
                        "(forward degrid + backward imaging of residuals), the samples sorted "
                        "once (lofar_plan_create) before the timed iterations",
            "value": n2 / (msG * 1e-3) / 1e6, "unit": "Mvis/s per iteration",
            "ms_per_iteration": msG, "chi2_first": h[0], "chi2_last": h[-1],
            "monotone": all(b <= a * (1 + 1e-6) for a, b in zip(h, h[1:]))}


def load_traffic(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(key)
    except ValueError:
        return None


def bench_stream(dev, n_base=16 * 1024 * 1024, nchan=16, nx=8192, nw=128, support=9, cycles=5):
    """Multi-channel streaming workload: host-resident observation streamed to the device,
    FP32C dirty image and error against the FP64 mode on the same samples, then
    ``cycles`` major cycles (degrid -> residual -> grid) on the resident samples."""
    import torch
    from lofar_amd import imager
    from lofar_amd.autograd import ImagingOperator, lipschitz
    from lofar_amd.stream import major_cycles, make_multichannel
    print("bench: generating %d x %d channel samples" % (n_base, nchan), file=sys.stderr,
          flush=True)
    obs = make_multichannel(n_base, nx, nchan=nchan, seed=2026, device=dev)
    s, secs, nbytes = obs.to_device(dev, chunk_samples=32 * 1024 * 1024)
    n = obs.num_samples
    res = 1.0 / int(obs.span)
    print("bench: streamed %.2f GB in %.1f ms" % (nbytes / 1e9, secs * 1e3), file=sys.stderr,
          flush=True)
    vis = (s["uu"], s["vv"], s["ww"], s["vis_real"], s["vis_img"], s["weight"])
    args = (nw, support, nx, nx, res, obs.wmin, obs.wmax)
    cube = torch.empty((nw, nx, nx), dtype=torch.complex64, device=dev)
    re32 = torch.empty((nx, nx), dtype=torch.float64, device=dev)
    im32 = torch.empty_like(re32)
    imager.image(*vis, *args, precision="fp32c", cube=cube, image_real=re32, image_imag=im32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    imager.image(*vis, *args, precision="fp32c", cube=cube, image_real=re32, image_imag=im32)
    torch.cuda.synchronize()
    ms_img = (time.perf_counter() - t0) * 1e3
    print("bench: stream fp32c image %.1f ms" % ms_img, file=sys.stderr, flush=True)
    img_stages = profile_stages(
        lambda: imager.image(*vis, *args, precision="fp32c", cube=cube, image_real=re32,
                             image_imag=im32), 1)
    # major cycles on the resident samples, reusing the FP32C cube as scratch
    op = ImagingOperator(s["uu"], s["vv"], s["ww"], None, nw, support, nx, nx, res, obs.wmin,
                         obs.wmax, precision="fp32c", cube=cube)
    step = 0.5 / lipschitz(op, s["weight"], iters=3)
    x0 = torch.zeros((nx, nx), dtype=torch.float64, device=dev)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    major_cycles(op, s["vis_real"], s["vis_img"], s["weight"], x0, x0.clone(), cycles=cycles,
                 step=step, history=hist)
    torch.cuda.synchronize()
    ms_cyc = (time.perf_counter() - t0) * 1e3 / cycles
    h = [float(v) for v in hist]
    cyc_stages = profile_stages(
        lambda: major_cycles(op, s["vis_real"], s["vis_img"], s["weight"], x0, x0.clone(),
                             cycles=1, step=step), 1)
    print("bench: stream major cycle %.1f ms, chi2 %.4e -> %.4e" % (ms_cyc, h[0], h[-1]),
          file=sys.stderr, flush=True)
    del op, cube, x0
    torch.cuda.empty_cache()
    # the same dirty image in FP64 mode (complex128 cube, 137 GB) for the error figure
    c64 = torch.empty((nw, nx, nx), dtype=torch.complex128, device=dev)
    r64, i64 = imager.image(*vis, *args, precision="fp64", cube=c64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    imager.image(*vis, *args, precision="fp64", cube=c64, image_real=r64, image_imag=i64)
    torch.cuda.synchronize()
    ms_img64 = (time.perf_counter() - t0) * 1e3
    rms, mx = rel_err(re32, im32, r64, i64)
    del c64
    torch.cuda.empty_cache()
    cube_bytes = nw * nx * nx * 8
    line = {
        "metric": "Mvis/s multi-channel streamed imaging and major cycles (FP32C)",
        "value": n / (ms_cyc * 1e-3) / 1e6, "unit": "Mvis/s per major cycle", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f32c", "data": "synthetic (seeded, lofar_amd.stream)",
        "config": {"workload": "%d samples (%d baseline samples x %d channels 120-168 MHz, uvw "
                               "scaled by nu/c), streamed from pinned host memory in 32M-sample "
                               "chunks, %dx%d grid, %d w-planes, support %d, complex64 cube"
                               % (n, n_base, nchan, nx, nx, nw, support),
                   "samples": n, "grid": [nx, nx], "w_planes": nw, "support": support},
        "stream_h2d_GBs": nbytes / secs / 1e9, "stream_ms": secs * 1e3,
        "image_ms": ms_img, "image_Mvis_s": n / (ms_img * 1e-3) / 1e6,
        "image_stages_ms": {k: round(v[0], 3) for k, v in img_stages.items()},
        "image_algorithmic_GBs": (36 * n + 4 * cube_bytes + 16 * nx * nx) / (ms_img * 1e-3) / 1e9,
        "major_cycle_ms": ms_cyc, "cycles": cycles, "chi2": h,
        "major_cycle_sort": "the samples sorted once (lofar_plan_create) before the timed cycles",
        "major_cycle_stages_ms": {k: round(v[0] * v[1], 3) for k, v in cyc_stages.items()},
        "image_rel_rms_err_vs_fp64": rms, "image_rel_max_err_vs_fp64": mx,
        "fp64_image_ms": ms_img64,
    }
    del s, vis, re32, im32, r64, i64
    torch.cuda.empty_cache()
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg (and with it the image errors against the "
                         "reference CPU path)")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the multi-channel streaming workload")
    ap.add_argument("--stream", action="store_true",
                    help="only the multi-channel streaming workload: 256M samples (16M "
                         "baseline samples x 16 channels, 120-168 MHz) streamed to the device "
                         "in 32M-sample chunks, 8192x8192 grid, 128 w-planes, support 9, FP32C, "
                         "5 major cycles; one JSON line of its own")
    ap.add_argument("--only", choices=("fp32c", "adjoint", "chi2", "dropin"),
                    help="only this workload (one JSON object of its own; for profiling one "
                         "workload at a time): fp32c = 64M samples FP32C imaging; adjoint, "
                         "chi2 and dropin on the 16M-sample headline samples")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (3 launches), the "
                         "known-size reference for scaling HBM byte counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: liblofar_hip has no multi-GPU path "
              "(the path would shard by w-plane with one image reduction)", file=sys.stderr)
        sys.exit(2)

    import torch
    from lofar_amd import imager
    from lofar_amd.synth import make_obs

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if a.stream:
        print(json.dumps(bench_stream(dev)), flush=True)
        return
    if a.calibrate_copy:
        src = torch.ones(2 ** 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 8,
                          "copy_bytes_written_per_launch": dst.numel() * 8}), flush=True)
        return
    nx = ny = 4096
    nw, support = 32, 7
    n2 = 16 * 1024 * 1024
    if a.only == "fp32c":
        print(json.dumps(bench_fp32c(dev, a, None, None, None)), flush=True)
        return
    print("bench: generating %d samples (%dx%d, %d planes)" % (n2, nx, ny, nw), file=sys.stderr, flush=True)
    d = make_obs(n2, nx, nw, seed=2024, backend="torch", device=dev)
    vis = [d[k] for k in FIELDS]
    if a.only:
        if a.only == "adjoint":
            cubeA = torch.empty((nw, ny, nx), dtype=torch.complex64, device=dev)
            res = bench_adjoint(d, vis, cubeA, n2, nx, ny, nw, support, a)
        elif a.only == "chi2":
            res = bench_chi2(d, n2, nx, ny, nw, support)
        else:
            res = bench_dropin(d, n2, nx, ny, nw, support)
        print(json.dumps(res), flush=True)
        return
    cube64 = torch.empty((nw, ny, nx), dtype=torch.complex128, device=dev)
    re64 = torch.empty((ny, nx), dtype=torch.float64, device=dev)
    im64 = torch.empty_like(re64)

    def step64():
        imager.image(*vis, nw, support, nx, ny, d["resolution"], d["wmin"], d["wmax"],
                     precision="fp64", cube=cube64, image_real=re64, image_imag=im64)

    ms64 = time_steps(step64, a.steps, a.warmup)
    value = n2 / (ms64 * 1e-3) / 1e6
    print("bench: fp64 %.3f ms/step  %.1f Mvis/s" % (ms64, value), file=sys.stderr, flush=True)
    prof = profile_stages(step64, 3)
    kb = per_kernel_bytes(n2, nx, ny, nw, 16)
    dom = max((k for k in prof if k in kb), key=lambda k: prof[k][0])
    dom_ms = prof[dom][0] / max(prof[dom][1], 1)
    achieved = kb[dom] / (dom_ms * 1e-3) / 1e9
    roofline = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": load_traffic("fp64_4096x4096x32_16M:" + dom),
                "traffic_source": TRAFFIC_SOURCE,
                "algorithmic_bytes": kb[dom], "avg_launch_ms": dom_ms,
                "stages_ms": {k: round(v[0], 4) for k, v in prof.items()},
                "end_to_end_GBs": kb["total"] / (ms64 * 1e-3) / 1e9,
                "gridder_lds_atomics": lds_atomics(n2, support, prof)}
    work = tempfile.mkdtemp(prefix="lofar_bench_")
    names = []
    if not a.no_cpu:
        save_workload(work, "fp64", d, nx, nw, support, re64, im64)
        names.append("fp64")

    extra = {}
    if not a.no_extra:
        del cube64
        torch.cuda.empty_cache()
        extra["fp32c"] = bench_fp32c(dev, a, value, work if not a.no_cpu else None, names)
        cubeA = torch.empty((nw, ny, nx), dtype=torch.complex64, device=dev)
        extra["adjoint"] = bench_adjoint(d, vis, cubeA, n2, nx, ny, nw, support, a)
        del cubeA
        torch.cuda.empty_cache()
        extra["chi2_fit"] = bench_chi2(d, n2, nx, ny, nw, support)
        torch.cuda.empty_cache()
        extra["dropin"] = bench_dropin(d, n2, nx, ny, nw, support)
        del d, vis, re64, im64
        torch.cuda.empty_cache()
        if not a.no_stream:
            extra["stream"] = bench_stream(dev)
            torch.cuda.empty_cache()

    cpu = None
    if names:
        print("bench: cpu_baseline leg (reference CPU path on this host)", file=sys.stderr,
              flush=True)
        res = cpu_baseline(work, names)
        wl = res.get("workloads", {})
        f64 = wl.get("fp64", {})
        cpu = dict(f64.get("baseline", {"error": res.get("error", "no fp64 workload")}))
        cpu["host"] = res.get("host")
        cpu["reference_path"] = f64.get("reference_path")
        if "fp32c" in wl and "fp32c" in extra:
            extra["fp32c"]["image_rel_rms_err_vs_reference"] = wl["fp32c"].get(
                "image_rel_rms_err_vs_reference")
            extra["fp32c"]["image_rel_max_err_vs_reference"] = wl["fp32c"].get(
                "image_rel_max_err_vs_reference")
    shutil.rmtree(work, ignore_errors=True)
    line = {
        "metric": METRIC, "value": value, "unit": "Mvis/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms64, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded, lofar_amd.synth)",
        "config": {"workload": "FP64 end-to-end imaging (grid+FFT+phase+plane sum), 16M samples, "
                               "4096x4096 grid, 32 w-planes, support 7, inputs resident in HBM",
                   "samples": n2, "grid": [nx, ny], "w_planes": nw, "support": support},
        "image_rel_rms_err_vs_reference": f64.get("image_rel_rms_err_vs_reference") if names else None,
        "image_rel_max_err_vs_reference": f64.get("image_rel_max_err_vs_reference") if names else None,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
