"""Writes tests/golden/*.npz: inputs and outputs of the reference's own hot-path C.

The reference's w-stacking_omp.c, w-stacking.cu (as C) and phase_correction.cu (as C,
PHASE_ON) are compiled in place by oracle/Makefile into oracle/_ref/; this script runs
them on small seeded cases and stores the results as fixtures, so the parity tests keep
a reference-produced answer where /root/reference is absent.  FFTW is not available in
this image, so plane transforms in the pipeline fixture use numpy.fft.ifft2, which has
FFTW_BACKWARD's sign and (with the 1/(NxNy) factor) fftw_data()'s scaling.

Run from the repository root after `make -C oracle`:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import RefLib  # noqa: E402


def case_inputs(seed, n, nx, ny, nw, F=1, P=1, edge=False, nan_frac=0.0, weights="ones"):
    rng = np.random.default_rng(seed)
    if edge:
        # positions that exercise the clipped windows at both grid edges and
        # slightly negative coordinates (truncation toward zero)
        u = np.concatenate([rng.uniform(-0.9, 4.0, n // 4), rng.uniform(nx - 4.0, nx + 2.0, n // 4),
                            rng.uniform(0, nx, n - 2 * (n // 4))]) / nx
        v = np.concatenate([rng.uniform(ny - 4.0, ny + 2.0, n // 4), rng.uniform(-0.9, 4.0, n // 4),
                            rng.uniform(0, ny, n - 2 * (n // 4))]) / nx
    else:
        u = rng.uniform(0.1, 0.9, n)
        v = rng.uniform(0.1, 0.9, n) * ny / nx
    w = rng.uniform(-0.9 / nw, 1.0 - 1e-9, n)
    vr = rng.standard_normal(n * F * P).astype(np.float32)
    vi = rng.standard_normal(n * F * P).astype(np.float32)
    if nan_frac:
        vr[rng.random(n * F * P) < nan_frac] = np.nan
    wt = np.ones(n * P, np.float32) if weights == "ones" else rng.uniform(0.1, 2.0, n * P).astype(np.float32)
    return dict(uu=u, vv=v, ww=w, vis_real=vr, vis_img=vi, weight=wt, nx=nx, ny=ny, nw=nw, F=F, P=P)


CASES = {
    "interior_s7": dict(seed=1, n=300, nx=32, ny=32, nw=3, support=7),
    "edges_s7": dict(seed=2, n=400, nx=32, ny=48, nw=2, support=7, edge=True),
    "nan_fp2_s5": dict(seed=3, n=200, nx=32, ny=32, nw=4, support=5, F=2, P=2, nan_frac=0.1,
                       weights="random"),
    "s9_nw1": dict(seed=4, n=250, nx=64, ny=32, nw=1, support=9, weights="random"),
    "even_s8": dict(seed=5, n=150, nx=16, ny=16, nw=2, support=8, edge=True),
}


def main():
    ref = RefLib(fast=False)
    if not ref.available:
        raise SystemExit("oracle/_ref is not built (make -C oracle with /root/reference present)")
    for name, c in CASES.items():
        support = c.pop("support")
        inp = case_inputs(**c)
        nx, ny, nw, F, P = inp["nx"], inp["ny"], inp["nw"], inp["F"], inp["P"]
        grid = ref.wstack(nw, inp["uu"], inp["vv"], inp["ww"], inp["vis_real"], inp["vis_img"],
                          inp["weight"], 1.0 / nx, 1.0 / nw, support, nx, ny,
                          freq_per_chan=F, polarizations=P, variant="serial")
        planes = np.fft.ifft2(grid, axes=(1, 2))
        res, wmin, wmax = 1.0 / (2 * nx), -37.5, 41.25
        img = ref.phase_correction(planes, res, wmin, wmax, num_threads=1)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), support=support, res=res, wmin=wmin,
                            wmax=wmax, grid=grid, image=img, **inp)
        print(name, grid.shape, float(np.abs(grid).max()), float(np.abs(img).max()))


if __name__ == "__main__":
    main()
