"""The reference CPU path at full workload sizes -- TEST INFRASTRUCTURE ONLY.

Used by tests/ (headline-size parity) and by bench.py's cpu_baseline leg
(oracle/cpu_baseline.py), never by the product.  It chains, on host arrays:

1. gridding: the reference's own w-stacking.cu OpenMP loop (w-stacking.cu:222-307,
   compiled unchanged as C into oracle/_ref/fast by oracle/Makefile), or its serial
   w-stacking_omp.c loop (w-stacking_omp.c:25-145); without oracle/_ref the clean-room
   restatement oracle_wstack (oracle.c, pinned bit-identical to the reference);
2. the per-plane FFTW_BACKWARD/(Nx*Ny) transform of fftw_data()
   (fourier_transform.c:31-61) with scipy.fft.ifft2, whose convention equals it
   (exp(+2 pi i ...) and the 1/(Nx*Ny) scale; tests/test_oracle.py pins scipy against the
   oracle's own DFT);
3. the reference's own phase_correction() OpenMP loop (phase_correction.cu:145-210,
   PHASE_ON), or without oracle/_ref the restatement oracle_phase_correction.

Every stage is timed with time.perf_counter (CLOCK_MONOTONIC), bucket names as the
reference's result.c:15-35 (kernel / fftw / phase).
"""
from __future__ import annotations

import os
import time

import numpy as np

from oracle import oracle as O


def host_threads():
    """CPU threads this process may use: the scheduler affinity, capped by
    OMP_NUM_THREADS when that is set (a shared host may preset it to one GPU's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_info():
    """nproc, the lscpu model name and the physical core count of this host."""
    model, sockets, cores_per_socket, threads_per_core = None, 1, None, 1
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Socket(s)" and v.isdigit():
                sockets = int(v)
            elif k.strip() == "Core(s) per socket" and v.isdigit():
                cores_per_socket = int(v)
            elif k.strip() == "Thread(s) per core" and v.isdigit():
                threads_per_core = int(v)
    except Exception:  # lscpu missing: /proc/cpuinfo only
        pass
    return {"nproc": os.cpu_count(), "affinity": host_threads(), "model": model,
            "physical_cores": (sockets * cores_per_socket) if cores_per_socket else None,
            "threads_per_core": threads_per_core}


def grid_reference(obs, nw, support, nx, ny, threads=None, variant="omp"):
    """The reference's wstack() on host arrays; returns (cube complex128 [nw,ny,nx],
    seconds, what ran)."""
    threads = threads or host_threads()
    cube = np.zeros((nw, ny, nx), np.complex128)
    ref = O.RefLib(fast=True)
    t0 = time.perf_counter()
    if ref.available:
        ref.wstack(nw, obs["uu"], obs["vv"], obs["ww"], obs["vis_real"], obs["vis_img"],
                   obs["weight"], 1.0 / nx, 1.0 / nw, support, nx, ny, grid=cube, variant=variant,
                   num_threads=threads)
        what = ("reference w-stacking.cu OpenMP gridder (%d threads)" % threads
                if variant == "omp" else "reference w-stacking_omp.c serial gridder")
    else:
        st, cube = O.wstack(nw, obs["uu"], obs["vv"], obs["ww"], obs["vis_real"],
                            obs["vis_img"], obs["weight"], 1.0 / nx, 1.0 / nw, support, nx, ny,
                            grid=cube)
        if st != O.OK:
            raise ValueError("oracle wstack failed: %d" % st)
        what = "oracle restatement of w-stacking_omp.c (serial)"
    return cube, time.perf_counter() - t0, what


def fft_reference(cube, threads=None):
    """fftw_data()'s per-plane FFTW_BACKWARD/(Nx*Ny), in place, plane by plane."""
    import scipy.fft
    threads = threads or host_threads()
    t0 = time.perf_counter()
    for w in range(cube.shape[0]):
        cube[w] = scipy.fft.ifft2(cube[w], workers=threads)
    return time.perf_counter() - t0


def phase_reference(cube, resolution, wmin, wmax, threads=None):
    """phase_correction() plane sum; returns (complex image, seconds, what ran)."""
    threads = threads or host_threads()
    ref = O.RefLib(fast=True)
    t0 = time.perf_counter()
    if ref.phase is not None:
        img = ref.phase_correction(cube, resolution, wmin, wmax, num_threads=threads)
        what = "reference phase_correction.cu OpenMP loop (%d threads)" % threads
    else:
        img = O.phase_correction(cube, resolution, wmin, wmax)
        what = "oracle restatement of phase_correction.cu (serial)"
    return img, time.perf_counter() - t0, what


def reference_image(obs, nw, support, nx, ny, resolution, wmin, wmax, threads=None):
    """Whole reference path: returns (complex image [ny,nx], {stage: seconds, ...})."""
    cube, t_grid, g_what = grid_reference(obs, nw, support, nx, ny, threads)
    t_fft = fft_reference(cube, threads)
    img, t_phase, p_what = phase_reference(cube, resolution, wmin, wmax, threads)
    del cube
    return img, {"kernel_s": t_grid, "fftw_s": t_fft, "phase_s": t_phase,
                 "gridder": g_what, "phase": p_what}


def image_errors(got, want):
    """Relative L2 error over the complex image and max |error| / max |want|."""
    d = got - want
    return (float(np.linalg.norm(d) / np.linalg.norm(want)),
            float(np.max(np.abs(d)) / np.max(np.abs(want))))
