"""CPU baseline and reference image errors -- TEST INFRASTRUCTURE, run by bench.py's
cpu_baseline leg as a child process (it never touches the GPU).

bench.py writes each benchmark workload's samples and the GPU's image to
``<work>/<name>.npz``; for every workload this script runs the reference CPU path
(oracle/refpath.py: the reference's own w-stacking.cu OpenMP gridder and
phase_correction.cu OpenMP loop, compiled unchanged at -O3 into oracle/_ref/fast, around
scipy's FFTW_BACKWARD/(Nx*Ny) plane transform) on the same samples and reports the GPU
image's relative RMS and max error against it.

For the first workload it also times the reference CPU path on this host, stage by stage
(bucket names of result.c:15-35):

* kernel: the OpenMP gridder (w-stacking.cu:222-307) on all N samples, threads = the
  cores this process may use; the serial gridder (w-stacking_omp.c:25-145, the
  reference's default build) on a 2M-sample part, reported beside it;
* fftw: one FFTW_ESTIMATE FFTW_BACKWARD plane transform, single-threaded, through MKL's
  FFTW3 interface (FFTW itself is not installed; the oracle's C FFT is the fallback), on
  2 planes, times Nw;
* phase: the OpenMP phase_correction loop (phase_correction.cu:145-210) on all planes.

    value = N / (t_kernel_omp + t_fft_plane * Nw + t_phase)   (Mvis/s)

Prints one JSON object.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import refpath  # noqa: E402
from oracle.oracle import RefLib, fft_planes  # noqa: E402

FIELDS = ("uu", "vv", "ww", "vis_real", "vis_img", "weight")


def load_mkl_fft():
    for path in ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.2",
                 "/opt/conda/lib/libmkl_rt.so.1"):
        if os.path.exists(path):
            try:
                L = ctypes.CDLL(path)
                L.fftw_plan_dft_2d.restype = ctypes.c_void_p
                L.fftw_plan_dft_2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_int, ctypes.c_uint]
                L.fftw_execute.argtypes = [ctypes.c_void_p]
                L.fftw_destroy_plan.argtypes = [ctypes.c_void_p]
                return L, path
            except (OSError, AttributeError):
                continue
    return None, None


def time_fft_plane(plane, reps):
    """Seconds per FFTW_BACKWARD plane transform (scaled by 1/(Nx*Ny)), single thread."""
    ny, nx = plane.shape
    L, mkl_path = load_mkl_fft()
    t = 0.0
    if L is not None:
        out = np.empty_like(plane)
        for _ in range(reps):
            plan = L.fftw_plan_dft_2d(ny, nx, plane.ctypes.data, out.ctypes.data, 1, 64)
            t0 = time.perf_counter()
            L.fftw_execute(plan)  # FFTW_BACKWARD (+1), FFTW_ESTIMATE (64)
            out *= 1.0 / (nx * ny)
            t += time.perf_counter() - t0
            L.fftw_destroy_plan(plan)
        kind = ("FFTW_ESTIMATE plan through MKL's FFTW3 interface (%s), MKL_NUM_THREADS=%s"
                % (mkl_path, os.environ.get("MKL_NUM_THREADS", "default")))
    else:
        for _ in range(reps):
            t0 = time.perf_counter()
            fft_planes(plane[None])
            t += time.perf_counter() - t0
        kind = "the oracle's C FFT (single thread)"
    return t / reps, kind


def run_workload(path, timed, grid_sample, fft_planes_n, threads):
    z = np.load(path)
    obs = {k: z[k] for k in FIELDS}
    nx, ny, nw, support = int(z["nx"]), int(z["ny"]), int(z["nw"]), int(z["support"])
    res, wmin, wmax = float(z["resolution"]), float(z["wmin"]), float(z["wmax"])
    got = z["image"]
    n = len(obs["uu"])
    out = {"samples": n, "grid": [nx, ny], "w_planes": nw, "support": support}
    cube, t_grid, g_what = refpath.grid_reference(obs, nw, support, nx, ny, threads, "omp")
    if timed:
        ref = RefLib(fast=True)
        m = min(grid_sample, n)
        part = {k: np.ascontiguousarray(obs[k][:m]) for k in FIELDS}
        scratch = np.zeros((nw, ny, nx), np.complex128)
        scratch.view(np.float64)[...] = 0.0  # pages faulted in before timing
        t0 = time.perf_counter()
        if ref.serial is not None:
            ref.wstack(nw, part["uu"], part["vv"], part["ww"], part["vis_real"],
                       part["vis_img"], part["weight"], 1.0 / nx, 1.0 / nw, support, nx, ny,
                       grid=scratch, variant="serial")
        t_serial = (time.perf_counter() - t0) / m
        del scratch
        t_fft1, fft_kind = time_fft_plane(np.ascontiguousarray(cube[nw // 2]), fft_planes_n)
    t_fft_mt = refpath.fft_reference(cube, threads)
    want, t_phase, p_what = refpath.phase_reference(cube, res, wmin, wmax, threads)
    del cube
    rms, mx = refpath.image_errors(got, want)
    out.update(image_rel_rms_err_vs_reference=rms, image_rel_max_err_vs_reference=mx,
               reference_path="%s; scipy.fft.ifft2 per plane (%d threads); %s"
                              % (g_what, threads, p_what),
               reference_s={"kernel": t_grid, "fftw": t_fft_mt, "phase": t_phase})
    if timed:
        total = t_grid + t_fft1 * nw + t_phase
        out["baseline"] = {
            "value": n / total / 1e6, "unit": "Mvis/s", "cores": threads, "kind": "reference",
            "sample": ("reference CPU path on this host: w-stacking.cu OpenMP gridder on all %d "
                       "samples (%d threads), %s on %d of the %d planes (x %d), "
                       "phase_correction.cu OpenMP loop on all planes (%d threads); "
                       "serial w-stacking_omp.c gridder timed on %d samples beside it; "
                       "-O3 -march=x86-64-v3 (the reference's Makefile builds at -O0)"
                       % (n, threads, fft_kind, fft_planes_n, nw, nw, threads, m)),
            "stage_s": {"kernel_omp": t_grid, "kernel_serial_per_mvis": t_serial * 1e6,
                        "kernel_serial_extrapolated": t_serial * n,
                        "fftw_per_plane_1thread": t_fft1, "phase_omp": t_phase,
                        "total": total},
            "serial_gridder_Mvis_s": n / (t_serial * n + t_fft1 * nw + t_phase) / 1e6,
        }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--work", required=True, help="directory of <name>.npz workloads")
    ap.add_argument("--names", default="fp64,fp32c", help="workloads; the first is timed")
    ap.add_argument("--grid-sample", type=int, default=2 * 1024 * 1024)
    ap.add_argument("--fft-planes", type=int, default=2)
    a = ap.parse_args()
    threads = refpath.host_threads()
    result = {"host": refpath.cpu_info(), "threads": threads, "workloads": {}}
    for k, name in enumerate(a.names.split(",")):
        path = os.path.join(a.work, name + ".npz")
        if not os.path.exists(path):
            continue
        t0 = time.perf_counter()
        result["workloads"][name] = run_workload(path, k == 0, a.grid_sample, a.fft_planes,
                                                 threads)
        print("cpu_baseline: %s done in %.1f s" % (name, time.perf_counter() - t0),
              file=sys.stderr, flush=True)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
