"""Parity of the HIP path (through the C ABI) with the oracle and the reference's
golden fixtures, at sizes the oracle finishes in seconds.

Tolerances (relative L2 over the whole array unless stated):
  FP64 mode: cube 1e-12, plane transform 1e-13, image 1e-10 (the FP64 parity bound);
  FP32C mode: cube 1e-6, plane transform 2e-6, image 1e-6.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_nan(a, b):
    """rel() over the finite pixels, after checking that both sides are NaN at the
    same pixels (l^2+m^2 > 1 gives sqrt of a negative number in the reference too)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    return rel(a[ok], b[ok])


def T(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def dev_inputs(d, dev):
    return [T(d[k], dev) for k in ("uu", "vv", "ww", "vis_real", "vis_img", "weight")]


def random_case(seed, n, nx, ny, nw, F=1, P=1, edge=True, nan=0.02):
    rng = np.random.default_rng(seed)
    if edge:
        u = rng.uniform(-0.99, nx + 3.0, n) / nx
        v = rng.uniform(-0.99, ny + 3.0, n) / nx
    else:
        u = rng.uniform(0.05, 0.95, n)
        v = rng.uniform(0.05, 0.95, n) * ny / nx
    w = rng.uniform(-0.99 / nw, 1 - 1e-12, n)
    vr = rng.standard_normal(n * F * P).astype(np.float32)
    vi = rng.standard_normal(n * F * P).astype(np.float32)
    if nan:
        vr[rng.random(n * F * P) < nan] = np.nan
    wt = rng.uniform(0.2, 2.0, n * P).astype(np.float32)
    return dict(uu=u, vv=v, ww=w, vis_real=vr, vis_img=vi, weight=wt)


def corner_case(seed, n, nx, nw):
    """Every sample within two cells of an interior corner of the 64x64 tiles (nx x nx
    grid): a 7-cell window then covers 4 tiles, so each record expands into 4 entries (and
    into 4 records when every tile is a super-tile)."""
    rng = np.random.default_rng(seed)
    t = nx // 64
    u = (64.0 * rng.integers(1, t, n) + rng.uniform(-2.0, 2.0, n)) / nx
    v = (64.0 * rng.integers(1, t, n) + rng.uniform(-2.0, 2.0, n)) / nx
    w = rng.uniform(0.0, 1 - 1e-12, n)
    vr = rng.standard_normal(n).astype(np.float32)
    vi = rng.standard_normal(n).astype(np.float32)
    wt = rng.uniform(0.2, 2.0, n).astype(np.float32)
    return dict(uu=u, vv=v, ww=w, vis_real=vr, vis_img=vi, weight=wt)


# ---------------------------------------------------------------------------
# gridding

@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_matches_reference_fixture(cuda, path, precision, tol):
    from lofar_amd import imager
    z = np.load(path)
    nx, ny, nw, F, P = (int(z[k]) for k in ("nx", "ny", "nw", "F", "P"))
    cube = imager.grid(*dev_inputs(z, cuda), nw, int(z["support"]), nx, ny, precision=precision,
                       freq_per_chan=F, polarizations=P)
    assert rel(H(cube), z["grid"]) < tol


@pytest.mark.parametrize("seed,n,nx,ny,nw,support,F,P,edge", [
    (1, 20000, 128, 128, 4, 7, 1, 1, False),
    (2, 20000, 96, 200, 3, 7, 1, 1, True),
    (3, 5000, 64, 64, 2, 9, 2, 2, True),
    (4, 5000, 130, 70, 5, 3, 1, 1, True),
    (5, 3000, 40, 40, 1, 1, 1, 1, True),
    (6, 3000, 64, 64, 3, 17, 1, 1, True),
    (7, 200000, 256, 256, 8, 7, 1, 1, False),
])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_matches_oracle(cuda, seed, n, nx, ny, nw, support, F, P, edge, precision, tol):
    from lofar_amd import imager
    d = random_case(seed, n, nx, ny, nw, F, P, edge)
    st, want = O.wstack(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / nx, 1 / nw, support, nx, ny, freq_per_chan=F, polarizations=P)
    assert st == O.OK
    cube = imager.grid(*dev_inputs(d, cuda), nw, support, nx, ny, precision=precision,
                       freq_per_chan=F, polarizations=P)
    assert rel(H(cube), want) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_is_bitwise_order_independent(cuda, precision, tol):
    """Tiles accumulate exact fixed-point sums: any sample order gives the same bits."""
    from lofar_amd import imager
    d = random_case(11, 60000, 192, 160, 4, F=2, P=1, edge=True)
    perm = np.random.default_rng(12).permutation(60000)
    dp = dict(d)
    for k in ("uu", "vv", "ww"):
        dp[k] = d[k][perm]
    for k in ("vis_real", "vis_img"):
        dp[k] = d[k].reshape(60000, 2)[perm].reshape(-1)
    dp["weight"] = d["weight"][perm]
    a = H(imager.grid(*dev_inputs(d, cuda), 4, 7, 192, 160, precision=precision,
                      freq_per_chan=2))
    b = H(imager.grid(*dev_inputs(dp, cuda), 4, 7, 192, 160, precision=precision,
                      freq_per_chan=2))
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    st, want = O.wstack(4, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / 192, 1 / 4, 7, 192, 160, freq_per_chan=2)
    assert st == O.OK and rel(a, want) < tol


@pytest.mark.parametrize("n,span", [(200000, 50), (20000, 50), (3000, 50), (900, 6)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_dense_tile_same_sign(cuda, precision, tol, n, span):
    """Samples packed inside one 64x64 tile (a span x span square of cells) with
    visibilities of one sign (positive real, negative imaginary parts), so that cell sums
    come close to the FP32C packed word's bound and every imaginary part carries into the
    real half: above 65536 entries, and with more than 256 entries touching a 4x4 block, the
    tile takes the two-word fixed point; below, the packed word; the cube matches the
    oracle either way."""
    from lofar_amd import imager
    rng = np.random.default_rng(61 + n)
    nx, nw = 128, 1
    d = {"uu": rng.uniform(70, 70 + span, n) / nx, "vv": rng.uniform(8, 8 + span, n) / nx,
         "ww": np.full(n, 0.25), "vis_real": rng.uniform(0.5, 1.5, n).astype(np.float32),
         "vis_img": -rng.uniform(0.5, 1.5, n).astype(np.float32),
         "weight": rng.uniform(0.5, 1.0, n).astype(np.float32)}
    got = H(imager.grid(*dev_inputs(d, cuda), nw, 7, nx, nx, precision=precision))
    st, want = O.wstack(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / nx, 1 / nw, 7, nx, nx)
    assert st == O.OK
    assert rel(got, want) < tol
    hot = np.abs(want) > 0.5 * np.abs(want).max()
    assert hot.sum() > 20 and rel(got[hot], want[hot]) < tol


@pytest.mark.parametrize("scale", [1e-30, 1e-8, 1e-3, 1e5, 1e20])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_is_scale_invariant(cuda, precision, tol, scale):
    """Visibilities scaled by 1e-30 .. 1e20 (model visibilities and chi^2 residuals are
    far from unit size): the fixed-point quanta follow the tiles' own magnitudes, so the
    cube matches the oracle at the mode's tolerance at every scale."""
    from lofar_amd import imager
    d = random_case(81, 30000, 192, 128, 3, edge=True)
    d["vis_real"] = (d["vis_real"].astype(np.float64) * scale).astype(np.float32)
    d["vis_img"] = (d["vis_img"].astype(np.float64) * scale).astype(np.float32)
    st, want = O.wstack(3, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / 192, 1 / 3, 7, 192, 128)
    assert st == O.OK
    got = H(imager.grid(*dev_inputs(d, cuda), 3, 7, 192, 128, precision=precision))
    assert rel(got, want) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_infinite_visibilities(cuda, precision, tol):
    """Infinite visibilities (real parts finite, so not skipped as NaN) propagate into their
    windows as in the reference: +inf, -inf, or NaN where both meet; the tiles holding them
    take the floating-point (ds_add_f64) path, the others the fixed-point one."""
    from lofar_amd import imager
    d = random_case(31, 20000, 256, 256, 3, edge=False, nan=0)
    vi = d["vis_img"].copy()
    vi[[17, 4000, 4001, 12345]] = [np.inf, -np.inf, np.inf, -np.inf]
    d["vis_img"] = vi
    got = H(imager.grid(*dev_inputs(d, cuda), 3, 7, 256, 256, precision=precision))
    st, want = O.wstack(3, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / 256, 1 / 3, 7, 256, 256)
    assert st == O.OK
    for part in (np.real, np.imag):
        a, b = part(got), part(want)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(np.isposinf(a), np.isposinf(b))
        assert np.array_equal(np.isneginf(a), np.isneginf(b))
    fin = np.isfinite(want.real) & np.isfinite(want.imag)
    assert fin.sum() > 0.9 * want.size
    assert rel(got[fin], want[fin]) < tol


@pytest.mark.parametrize("grid_cfg", ["1", "2"])
def test_grid_accumulation_variants(cuda, grid_cfg):
    """The other accumulation paths (LOFAR_GRID_CFG, read once per process: 1 = LDS
    ds_add_f64 in both modes instead of exact fixed point, 2 = FP32C fixed point with two
    64-bit words per cell instead of one packed word) give the oracle's cube, in a child
    process with the setting applied."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, LOFAR_GRID_CFG=grid_cfg)
    r = subprocess.run([sys.executable, "-c", _BUCKET_SCRIPT, root], env=env, cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "BUCKETS_OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


_BUCKET_SCRIPT = r"""
import os, sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
from lofar_amd import imager
from tests.test_gpu_parity import corner_case, random_case, dev_inputs, rel, H
dev = torch.device("cuda:0")
for seed, n, nx, ny, nw in ((21, 40000, 300, 200, 3), (22, 3000, 1100, 64, 2),
                            (23, 30000, 512, 512, 2)):
    d = random_case(seed, n, nx, ny, nw) if seed != 23 else corner_case(seed, n, nx, nw)
    st, want = O.wstack(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / nx, 1 / nw, 7, nx, ny)
    assert st == O.OK
    for prec, tol in (("fp64", 1e-12), ("fp32c", 1e-6)):
        e = rel(H(imager.grid(*dev_inputs(d, dev), nw, 7, nx, ny, precision=prec)), want)
        assert e < tol, (seed, prec, e)
    uu, vv, ww, _, _, wt = dev_inputs(d, dev)
    mr, mi = imager.degrid_cube(torch.from_numpy(want.copy()).to(dev), uu, vv, ww, wt, 7)
    st, model = O.degrid(want, d["uu"], d["vv"], d["ww"], d["weight"], 1 / nx, 1 / nw, 7)
    assert st == O.OK and rel(H(mr) + 1j * H(mi), model) < 1e-12
    # a plan's sort (sample indices in the records) under the same setting: planned images
    # carry the unplanned ones' bits (to rounding where the gridder adds doubles in LDS,
    # LOFAR_GRID_CFG=1, whose order is not fixed)
    exact = os.environ.get("LOFAR_GRID_CFG", "0") != "1"
    if nx == ny:
        uu, vv, ww, vr, vi, wt = dev_inputs(d, dev)
        for prec in ("fp64", "fp32c"):
            plan = imager.Plan(uu, vv, ww, wt, nw, 7, nx, ny, precision=prec)
            a = plan.image(vr, vi, wt, 0.5 / nx, -10.0, 10.0)
            b = imager.image(uu, vv, ww, vr, vi, wt, nw, 7, nx, ny, 0.5 / nx, -10.0, 10.0,
                             precision=prec)
            for x, y in zip(a, b):
                if exact:
                    assert np.array_equal(H(x).view(np.uint8), H(y).view(np.uint8)), (seed, prec)
                else:
                    assert rel(H(x), H(y)) < (1e-12 if prec == "fp64" else 1e-6), (seed, prec)
            plan.close()
print("BUCKETS_OK")
"""


@pytest.mark.parametrize("cfg", ["1,256,256", "2,1000,300", "16,300,100000", "8,1000000,8192",
                                 "1,100000,256", "global", "stage=0", "stage=1", "stage=2",
                                 "order=0", "order=4", "nt=256", "nt=1024"])
def test_bucket_sort_configurations(cuda, cfg):
    """Other super-tile sides, level-1 chunk and level-2 block sizes (LOFAR_BUCKET_CFG, read once per
    process) give the same cube; "global": level 1 with global histograms and cursors (the
    path for more super-tiles than the LDS histograms hold, forced here with
    LOFAR_L1_LDS_MAX=2); "stage=b": LOFAR_BUCKET_STAGE=b (0: no scatter staged through LDS,
    1: level 1 only, 2: level 2 only); "order=o": LOFAR_TILE_ORDER=o (launch order of the
    per-tile kernels); "nt=t": LOFAR_GRID_NT=t (gridder workgroup size in both modes).
    "1,100000,256" overflows the staged level 1's round buffer (corner samples: 4 records
    each).  Run in a child process with the setting applied."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if cfg == "global":
        env = dict(os.environ, LOFAR_L1_LDS_MAX="2")
    elif cfg.startswith("stage="):
        env = dict(os.environ, LOFAR_BUCKET_STAGE=cfg[6:])
    elif cfg.startswith("order="):
        env = dict(os.environ, LOFAR_TILE_ORDER=cfg[6:])
    elif cfg.startswith("nt="):
        env = dict(os.environ, LOFAR_GRID_NT=cfg[3:])
    else:
        env = dict(os.environ, LOFAR_BUCKET_CFG=cfg)
    r = subprocess.run([sys.executable, "-c", _BUCKET_SCRIPT, root], env=env, cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "BUCKETS_OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_and_degrid_tile_corner_samples(cuda, precision, tol):
    """Samples at tile corners (4 entries per record: more than a staged level-2 round's
    LDS buffer holds, the rest stored directly) grid to the oracle's cube; their
    degridding (home sort) matches the oracle's."""
    import torch
    from lofar_amd import imager
    nx, nw = 512, 2
    d = corner_case(24, 30000, nx, nw)
    st, want = O.wstack(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / nx, 1 / nw, 7, nx, nx)
    assert st == O.OK
    assert rel(H(imager.grid(*dev_inputs(d, cuda), nw, 7, nx, nx, precision=precision)), want) < tol
    uu, vv, ww, _, _, wt = dev_inputs(d, cuda)
    mr, mi = imager.degrid_cube(torch.from_numpy(want.copy()).to(cuda), uu, vv, ww, wt, 7)
    st, model = O.degrid(want, d["uu"], d["vv"], d["ww"], d["weight"], 1 / nx, 1 / nw, 7)
    assert st == O.OK and rel(H(mr) + 1j * H(mi), model) < 1e-12


def test_grid_accumulate_and_empty(cuda):
    import torch
    from lofar_amd import imager
    d = random_case(8, 4000, 64, 64, 2)
    c1 = imager.grid(*dev_inputs(d, cuda), 2, 7, 64, 64)
    c2 = imager.grid(*dev_inputs(d, cuda), 2, 7, 64, 64, cube=c1.clone(), accumulate=True)
    assert rel(H(c2), 2 * H(c1)) < 1e-14
    empty = [t[:0] for t in dev_inputs(d, cuda)]
    c3 = imager.grid(*empty, 2, 7, 64, 64, cube=torch.full_like(c1, 3.0))
    assert not H(c3).any()


def test_grid_rejects_bad_w_and_writes_nothing(cuda):
    import torch
    from lofar_amd import imager
    from lofar_amd._lib import ECOORD, LofarError
    d = random_case(9, 1000, 32, 32, 2)
    d["ww"][17] = 1.0  # (int)(1.0/0.5) = 2 = Nw: outside the cube
    cube = torch.full((2, 32, 32), 5.0, dtype=torch.complex128, device=cuda)
    with pytest.raises(LofarError) as e:
        imager.grid(*dev_inputs(d, cuda), 2, 7, 32, 32, cube=cube)
    assert e.value.code == ECOORD
    torch.cuda.synchronize()
    assert (H(cube) == 5.0).all()


def test_wstack_dropin_accumulates_like_reference(cuda):
    from lofar_amd import imager
    z = np.load(os.path.join(GOLDEN, "nan_fp2_s5.npz"))
    nx, ny, nw, F, P = (int(z[k]) for k in ("nx", "ny", "nw", "F", "P"))
    n = len(z["uu"])
    start = np.random.default_rng(0).standard_normal(2 * nw * nx * ny)
    grid = start.copy()
    imager.wstack(nw, n, F, P, z["uu"], z["vv"], z["ww"], z["vis_real"], z["vis_img"], z["weight"],
                  1 / nx, 1 / nw, int(z["support"]), nx, ny, grid)
    want = start + z["grid"].view(np.float64).ravel()
    assert rel(grid, want) < 1e-12


def test_wstack_dropin_error_leaves_grid(cuda):
    from lofar_amd import imager
    from lofar_amd._lib import ECOORD, LofarError
    d = random_case(10, 100, 16, 16, 2)
    d["ww"][3] = float("nan")
    grid = np.full(2 * 2 * 16 * 16, 7.0)
    with pytest.raises(LofarError) as e:
        imager.wstack(2, 100, 1, 1, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"],
                      d["weight"], 1 / 16, 1 / 2, 7, 16, 16, grid)
    assert e.value.code == ECOORD and (grid == 7.0).all()


# ---------------------------------------------------------------------------
# plane transform

@pytest.mark.parametrize("nx,ny,nw", [(2, 2, 3), (4, 16, 2), (16, 16, 2), (32, 8, 1), (64, 64, 3),
                                      (128, 512, 1), (256, 256, 2), (1024, 1024, 1), (1024, 1024, 4),
                                      (2048, 64, 1), (64, 2048, 1), (4096, 4096, 1)])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_matches_numpy(cuda, nx, ny, nw, sign):
    from lofar_amd import imager
    rng = np.random.default_rng(nx * 7 + ny + nw)
    x = rng.standard_normal((nw, ny, nx)) + 1j * rng.standard_normal((nw, ny, nx))
    want = np.fft.ifft2(x, axes=(1, 2)) if sign > 0 else np.fft.fft2(x, axes=(1, 2)) / (nx * ny)
    got64 = H(imager.fft(T(x, cuda), sign=sign))
    assert rel(got64, want) < 1e-13
    got32 = H(imager.fft(T(x.astype(np.complex64), cuda), sign=sign))
    assert rel(got32, want) < 2e-6


def test_fft_8192_single_plane(cuda):
    from lofar_amd import imager
    rng = np.random.default_rng(81)
    x = (rng.standard_normal((1, 8192, 8192)) + 1j * rng.standard_normal((1, 8192, 8192))).astype(np.complex64)
    want = np.fft.ifft2(x.astype(np.complex128), axes=(1, 2))
    assert rel(H(imager.fft(T(x, cuda))), want) < 2e-6
    assert rel(H(imager.fft(T(x.astype(np.complex128), cuda))), want) < 1e-13


def test_fft_in_place_and_rejects_unsupported_sides(cuda):
    """In-place transforms; a side with a prime factor above 13 (17, 8191) is rejected."""
    from lofar_amd import imager
    from lofar_amd._lib import ESIZE, LofarError
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 64, 32)) + 1j * rng.standard_normal((2, 64, 32))
    t = T(x, cuda)
    imager.fft(t, out=t)
    assert rel(H(t), np.fft.ifft2(x, axes=(1, 2))) < 1e-13
    for shape in ((1, 12, 17), (1, 8191, 4)):
        with pytest.raises(LofarError) as e:
            imager.fft(T(np.zeros(shape, complex), cuda))
        assert e.value.code == ESIZE


@pytest.mark.parametrize("nx,ny,nw", [(12, 16, 2), (96, 60, 3), (300, 300, 2), (1000, 1000, 1),
                                      (6, 10, 1), (3000, 8, 1), (8, 3000, 1), (7, 11, 2),
                                      (8000, 4, 1), (13, 26, 1), (6000, 20, 1)])
@pytest.mark.parametrize("sign", [1, -1])
def test_fft_mixed_radix_matches_numpy(cuda, nx, ny, nw, sign):
    """Sides that are not powers of two (fftw_data plans any size, fourier_transform.c:25-27)
    take the mixed-radix transforms: FP64 within 1e-13 and FP32 within 2e-6 of numpy."""
    from lofar_amd import imager
    rng = np.random.default_rng(nx * 7 + ny)
    x = rng.standard_normal((nw, ny, nx)) + 1j * rng.standard_normal((nw, ny, nx))
    want = np.fft.ifft2(x, axes=(1, 2)) if sign > 0 else np.fft.fft2(x, axes=(1, 2)) / (nx * ny)
    got64 = H(imager.fft(T(x, cuda), sign=sign))
    assert rel(got64, want) < 1e-13
    got32 = H(imager.fft(T(x.astype(np.complex64), cuda), sign=sign))
    assert rel(got32, want) < 2e-6


@pytest.mark.parametrize("nx,ny,nw", [(300, 300, 4), (96, 160, 3)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-6)])
def test_mixed_radix_image_and_adjoint_match_oracle(cuda, nx, ny, nw, precision, tol):
    """Imaging and its adjoint on grids whose sides are not powers of two (and non-square)
    against the oracle, and the drop-in fftw_data on such a cube."""
    from lofar_amd import imager
    d = random_case(nx + ny, 30000, nx, ny, nw, edge=False, nan=0.01)
    res, wmin, wmax = 1.0 / (2 * max(nx, ny)), -50.0, 70.0
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7,
                   nx, ny, res, wmin, wmax)
    r, i = imager.image(*dev_inputs(d, cuda), nw, 7, nx, ny, res, wmin, wmax,
                        precision=precision)
    assert rel_nan(H(r) + 1j * H(i), want) < tol
    rng = np.random.default_rng(nx)
    y = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    want_m = O.degrid_image(y, nw, d["uu"], d["vv"], d["ww"], d["weight"], 7, res, wmin, wmax)
    mr, mi = imager.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), nw, 7, res,
                           wmin, wmax, precision=precision)
    assert rel(H(mr) + 1j * H(mi), want_m) < (1e-10 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        grid = np.zeros(2 * nw * nx * ny)
        imager.wstack(nw, len(d["uu"]), 1, 1, d["uu"], d["vv"], d["ww"], d["vis_real"],
                      d["vis_img"], d["weight"], 1 / nx, 1 / nw, 7, nx, ny, grid)
        gridss = np.zeros_like(grid)
        imager.fftw_data(grid, gridss, nx, ny, nw)
        cube = grid.view(np.complex128).reshape(nw, ny, nx)
        assert rel(gridss.view(np.complex128).reshape(nw, ny, nx),
                   np.fft.ifft2(cube, axes=(1, 2))) < 1e-13


def test_fftw_data_dropin(cuda):
    from lofar_amd import imager
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, 128, 64)) + 1j * rng.standard_normal((3, 128, 64))
    out = np.zeros(2 * x.size)
    imager.fftw_data(x.view(np.float64).ravel(), out, 64, 128, 3)
    assert rel(out.view(np.complex128).reshape(x.shape), O.fft_planes(x)) < 1e-13


# ---------------------------------------------------------------------------
# phase correction

@pytest.mark.parametrize("nx,ny,nw,wmin,wmax", [(64, 64, 8, -300.0, 400.0), (128, 32, 1, -5.0, 5.0),
                                                (32, 96, 5, 0.0, 1200.0)])
def test_phase_dropin_matches_oracle(cuda, nx, ny, nw, wmin, wmax):
    from lofar_amd import imager
    rng = np.random.default_rng(nx + ny + nw)
    cube = rng.standard_normal((nw, ny, nx)) + 1j * rng.standard_normal((nw, ny, nx))
    res = 1 / (2.0 * nx)
    re0 = rng.standard_normal(nx * ny)
    im0 = rng.standard_normal(nx * ny)
    re, im = re0.copy(), im0.copy()
    imager.phase_correction(cube.view(np.float64).ravel(), re, im, nx, ny, nw, nx, ny, res, wmin, wmax)
    want = O.phase_correction(cube, res, wmin, wmax, image=(re0 + 1j * im0).reshape(ny, nx))
    got = (re + 1j * im).reshape(ny, nx)
    assert rel_nan(got, want) < 1e-12
    # device entry point, both precisions, overwrite semantics
    r, i = imager.phase(T(cube, cuda), res, wmin, wmax)
    assert rel_nan(H(r) + 1j * H(i), O.phase_correction(cube, res, wmin, wmax)) < 1e-12
    r, i = imager.phase(T(cube.astype(np.complex64), cuda), res, wmin, wmax)
    want32 = O.phase_correction(cube.astype(np.complex64), res, wmin, wmax)
    assert rel_nan(H(r) + 1j * H(i), want32) < 1e-12


def test_phase_uses_global_axes_for_coordinates(cuda):
    # xaxistot/yaxistot drive the coordinate shift (phase_correction.cu:164-169)
    from lofar_amd import imager
    rng = np.random.default_rng(6)
    nx, ny, nw = 32, 16, 3
    cube = rng.standard_normal((nw, ny, nx)) + 1j * rng.standard_normal((nw, ny, nx))
    re, im = np.zeros(nx * ny), np.zeros(nx * ny)
    imager.phase_correction(cube.view(np.float64).ravel(), re, im, nx, ny, nw, nx, 64, 1 / 64, -9.0, 11.0)
    want = np.zeros((ny, nx), complex)
    lo = O.lib()
    wr, wi = np.zeros(nx * ny), np.zeros(nx * ny)
    lo.oracle_phase_correction(O._d(cube.view(np.float64).ravel()), O._d(wr), O._d(wi), nx, ny, nw,
                               nx, 64, 1 / 64, -9.0, 11.0, 1)
    want = (wr + 1j * wi).reshape(ny, nx)
    assert rel((re + 1j * im).reshape(ny, nx), want) < 1e-12


def test_phase_adjoint_matches_oracle(cuda):
    from lofar_amd import imager
    rng = np.random.default_rng(7)
    ny, nx, nw = 48, 64, 6
    img = rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))
    want = O.phase_adjoint(img, nw, 1 / 128, -70.0, 90.0)
    got = imager.phase_adjoint(T(img.real.copy(), cuda), T(img.imag.copy(), cuda), nw, 1 / 128, -70.0, 90.0)
    assert rel(H(got), want) < 1e-12


# ---------------------------------------------------------------------------
# whole hot path and its adjoint

def synth(n, nx, nw, seed):
    from lofar_amd.synth import make_obs
    return make_obs(n, nx, nw, seed=seed)


@pytest.mark.parametrize("n,nx,nw,support", [(30000, 256, 4, 7), (1 << 20, 1024, 8, 7),
                                             (50000, 512, 16, 9), (30000, 256, 4, 13),
                                             (30000, 256, 4, 17)])
def test_image_pipeline_matches_oracle(cuda, n, nx, nw, support):
    from lofar_amd import imager
    d = synth(n, nx, nw, seed=n + nx)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], support,
                   nx, nx, d["resolution"], d["wmin"], d["wmax"])
    r, i = imager.image(*dev_inputs(d, cuda), nw, support, nx, nx, d["resolution"], d["wmin"],
                        d["wmax"], precision="fp64")
    e64 = rel(H(r) + 1j * H(i), want)
    r, i = imager.image(*dev_inputs(d, cuda), nw, support, nx, nx, d["resolution"], d["wmin"],
                        d["wmax"], precision="fp32c")
    e32 = rel(H(r) + 1j * H(i), want)
    print("image rel err fp64 %.3e fp32c %.3e" % (e64, e32))
    assert e64 < 1e-10
    assert e32 < 1e-6


@pytest.mark.parametrize("n,nx,nw", [(3000, 512, 8), (40000, 2048, 2)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-6)])
def test_image_ignores_stale_scratch_cube(cuda, precision, tol, n, nx, nw):
    """The fused path leaves tiles without samples unwritten (the column transform reads
    them as zeros) and whole tile columns without samples unwritten after the column
    transform (the row pass reads them as zeros): stale NaNs in the caller's scratch cube
    must not leak into the image.  Sparse coverage leaves most tiles of most planes empty;
    at 2048 the disk of samples leaves edge tile columns empty, one of them paired with a
    column holding samples in a complex-float prefetch run."""
    import torch
    from lofar_amd import imager
    d = synth(n, nx, nw, seed=77)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7,
                   nx, nx, d["resolution"], d["wmin"], d["wmax"])
    cube = torch.full((nw, nx, nx), float("nan"), dtype=imager.cube_dtype(precision),
                      device=cuda)
    for _ in range(2):
        r, i = imager.image(*dev_inputs(d, cuda), nw, 7, nx, nx, d["resolution"], d["wmin"],
                            d["wmax"], precision=precision, cube=cube)
        assert rel(H(r) + 1j * H(i), want) < tol


def test_image_8192_split_row_pass(cuda):
    """8192-wide grids run the row transform and the phase/plane sum as two passes in both
    modes (the fused row pass does not fit its 1024-thread workgroups): the FP32C image
    agrees with the FP64 one on the same samples, and accumulate mode adds to the image."""
    from lofar_amd import imager
    nx, nw = 8192, 3
    d = synth(200000, nx, nw, seed=8192)
    args = (nw, 7, nx, nx, d["resolution"], d["wmin"], d["wmax"])
    r64, i64 = imager.image(*dev_inputs(d, cuda), *args, precision="fp64")
    want = H(r64) + 1j * H(i64)
    r32, i32 = imager.image(*dev_inputs(d, cuda), *args, precision="fp32c")
    e = rel(H(r32) + 1j * H(i32), want)
    print("8192 fp32c vs fp64 %.3e" % e)
    assert e < 1e-6
    imager.image(*dev_inputs(d, cuda), *args, precision="fp32c", image_real=r32, image_imag=i32,
                 accumulate=True)
    assert rel(H(r32) + 1j * H(i32), 2 * want) < 1e-6


_LAYOUT_SCRIPT = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
from oracle import oracle as O
from lofar_amd import imager
from lofar_amd.synth import make_obs
from tests.test_gpu_parity import dev_inputs, rel, H
dev = torch.device("cuda:0")
for n, nx, nw in ((40000, 256, 4), (5000, 64, 3)):
    d = make_obs(n, nx, nw, seed=nx + nw)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7,
                   nx, nx, d["resolution"], d["wmin"], d["wmax"])
    for prec, tol in (("fp64", 1e-10), ("fp32c", 1e-6)):
        cube = torch.full((nw, nx, nx), float("nan"), dtype=imager.cube_dtype(prec), device=dev)
        r, i = imager.image(*dev_inputs(d, dev), nw, 7, nx, nx, d["resolution"], d["wmin"],
                            d["wmax"], precision=prec, cube=cube)
        e = rel(H(r) + 1j * H(i), want)
        assert e < tol, (n, nx, prec, e)
print("LAYOUT_OK")
"""


@pytest.mark.parametrize("sm_log", ["-1", "2", "3", "4", "6"])
def test_image_scratch_layouts(cuda, sm_log):
    """The fused path's scratch cube in row-major or strip-major planes of 2^sm_log
    columns (LOFAR_SM_LOG, read once per process) gives the same image."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, LOFAR_SM_LOG=sm_log)
    r = subprocess.run([sys.executable, "-c", _LAYOUT_SCRIPT, root], env=env, cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "LAYOUT_OK" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]


def test_dropin_chain_matches_oracle(cuda):
    # wstack -> fftw_data -> phase_correction through the host drop-ins, as gridding.c,
    # fourier_transform.c:7-74 and :173 call them
    from lofar_amd import imager
    d = synth(40000, 256, 8, seed=11)
    nx = ny = 256
    nw = 8
    grid = np.zeros(2 * nw * nx * ny)
    imager.wstack(nw, len(d["uu"]), 1, 1, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"],
                  d["weight"], 1 / nx, 1 / nw, 7, nx, ny, grid)
    gridss = np.zeros_like(grid)
    imager.fftw_data(grid, gridss, nx, ny, nw)
    re, im = np.zeros(nx * ny), np.zeros(nx * ny)
    imager.phase_correction(gridss, re, im, nx, ny, nw, nx, ny, d["resolution"], d["wmin"], d["wmax"])
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7, nx, ny,
                   d["resolution"], d["wmin"], d["wmax"])
    assert rel((re + 1j * im).reshape(ny, nx), want) < 1e-10


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-5)])
def test_degrid_cube_matches_oracle(cuda, precision, tol):
    from lofar_amd import imager
    d = random_case(12, 30000, 128, 96, 4, edge=True, nan=0)
    rng = np.random.default_rng(13)
    cube = rng.standard_normal((4, 96, 128)) + 1j * rng.standard_normal((4, 96, 128))
    st, want = O.degrid(cube, d["uu"], d["vv"], d["ww"], d["weight"], 1 / 128, 1 / 4, 7)
    c = T(cube if precision == "fp64" else cube.astype(np.complex64), cuda)
    mr, mi = imager.degrid_cube(c, T(d["uu"], cuda), T(d["vv"], cuda), T(d["ww"], cuda),
                                T(d["weight"], cuda), 7)
    assert st == O.OK and rel(H(mr) + 1j * H(mi), want) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-5)])
def test_degrid_matches_oracle_adjoint(cuda, precision, tol):
    from lofar_amd import imager
    d = synth(20000, 256, 8, seed=14)
    rng = np.random.default_rng(15)
    img = rng.standard_normal((256, 256)) + 1j * rng.standard_normal((256, 256))
    want = O.degrid_image(img, 8, d["uu"], d["vv"], d["ww"], d["weight"], 7, d["resolution"],
                          d["wmin"], d["wmax"])
    mr, mi = imager.degrid(T(img.real.copy(), cuda), T(img.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), 8, 7,
                           d["resolution"], d["wmin"], d["wmax"], precision=precision)
    assert rel(H(mr) + 1j * H(mi), want) < tol


@pytest.mark.parametrize("support", [9, 13, 17])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-5)])
def test_degrid_wide_tile_windows_match_oracle(cuda, precision, tol, support):
    """The tile degridder's 9- and 17-cell window kernels (halos of 4 and 8 cells staged in
    LDS, every region load issued before the LDS stores) against the oracle's interpolation,
    samples reaching every grid edge."""
    from lofar_amd import imager
    nx, nw = 256, 4
    d = random_case(16, 20000, nx, nx, nw, edge=True, nan=0)
    res, wmin, wmax = 0.5 / nx, -30.0, 30.0  # every pixel inside l^2 + m^2 < 1
    rng = np.random.default_rng(17)
    img = rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx))
    want = O.degrid_image(img, nw, d["uu"], d["vv"], d["ww"], d["weight"], support, res, wmin,
                          wmax)
    mr, mi = imager.degrid(T(img.real.copy(), cuda), T(img.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), nw, support,
                           res, wmin, wmax, precision=precision)
    got = H(mr).astype(np.float64) + 1j * H(mi).astype(np.float64)
    assert rel(got, want) < tol


@pytest.mark.parametrize("nx,nw", [(256, 8), (8192, 1)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-5)])
def test_degrid_stores_only_the_tiles_it_reads(cuda, precision, tol, nx, nw):
    """The adjoint's last column transform stores only the tiles holding samples and their
    neighbours (the interpolation's halo), and skips tile columns without any: with a
    NaN-filled scratch cube and samples in a few clusters (most tiles never stored), the
    degridding still matches the oracle, and the planned degridding gives the same bits."""
    import torch
    from lofar_amd import imager
    rng = np.random.default_rng(16)
    n = 3000
    cu = rng.choice([0.1, 0.45, 0.8], n) + rng.normal(0, 0.02, n)
    cv = rng.choice([0.2, 0.7], n) + rng.normal(0, 0.02, n)
    d = {"uu": np.clip(cu, 0.0, 0.999), "vv": np.clip(cv, 0.0, 0.999),
         "ww": rng.uniform(0, 1 - 1e-9, n), "weight": rng.uniform(0.5, 1, n).astype(np.float32)}
    img = rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx))
    res, wmin, wmax = 0.5 / nx, -20.0, 20.0
    uu, vv, ww, wt = (T(d[k], cuda) for k in ("uu", "vv", "ww", "weight"))
    cube = torch.full((nw, nx, nx), float("nan"), dtype=imager.cube_dtype(imager._prec(precision)),
                      device=cuda)
    mr, mi = imager.degrid(T(img.real.copy(), cuda), T(img.imag.copy(), cuda), uu, vv, ww, wt,
                           nw, 7, res, wmin, wmax, precision=precision, cube=cube)
    got = H(mr).astype(np.float64) + 1j * H(mi).astype(np.float64)
    assert np.isfinite(got).all()
    if nx <= 1024:
        want = O.degrid_image(img, nw, d["uu"], d["vv"], d["ww"], d["weight"], 7, res, wmin, wmax)
        assert rel(got, want) < tol
    plan = imager.Plan(uu, vv, ww, wt, nw, 7, nx, nx, precision=precision)
    cube.fill_(float("nan"))
    pr, pi = plan.degrid(T(img.real.copy(), cuda), T(img.imag.copy(), cuda), res, wmin, wmax,
                         cube=cube)
    assert np.array_equal(H(pr), H(mr)) and np.array_equal(H(pi), H(mi))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_adjoint_dot_test_8192(cuda, precision, tol):
    """8192-wide grids run the adjoint's conjugate screen and row transform as two passes:
    the dot-product identity against the imaging path still holds."""
    from lofar_amd import imager
    nx, nw = 8192, 2
    d = synth(50000, nx, nw, seed=19)
    x = d["vis_real"].astype(np.float64) + 1j * d["vis_img"].astype(np.float64)
    r, i = imager.image(*dev_inputs(d, cuda), nw, 7, nx, nx, d["resolution"], d["wmin"],
                        d["wmax"], precision=precision)
    ax = H(r) + 1j * H(i)
    rng = np.random.default_rng(20)
    # y = Ax (normalised) plus noise: <y, Ax> is then not a zero-mean sum whose
    # occasional small values would inflate the relative error of a correct adjoint
    y = ax / np.sqrt(np.mean(np.abs(ax) ** 2)) + (rng.standard_normal((nx, nx)) +
                                                   1j * rng.standard_normal((nx, nx)))
    mr, mi = imager.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), nw, 7,
                           d["resolution"], d["wmin"], d["wmax"], precision=precision)
    ahy = H(mr).astype(np.float64) + 1j * H(mi).astype(np.float64)
    lhs, rhs = np.vdot(y, ax), np.vdot(ahy, x)
    print("dot test 8192 %s: %.3e" % (precision, abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs - rhs) < tol * abs(lhs)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_adjoint_dot_test(cuda, precision, tol):
    import torch
    from lofar_amd import imager
    d = synth(100000, 512, 8, seed=16)
    x = d["vis_real"].astype(np.float64) + 1j * d["vis_img"].astype(np.float64)
    r, i = imager.image(*dev_inputs(d, cuda), 8, 7, 512, 512, d["resolution"], d["wmin"], d["wmax"],
                        precision=precision)
    ax = H(r) + 1j * H(i)
    rng = np.random.default_rng(17)
    y = rng.standard_normal((512, 512)) + 1j * rng.standard_normal((512, 512))
    mr, mi = imager.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), 8, 7,
                           d["resolution"], d["wmin"], d["wmax"], precision=precision)
    ahy = H(mr).astype(np.float64) + 1j * H(mi).astype(np.float64)
    lhs, rhs = np.vdot(y, ax), np.vdot(ahy, x)
    print("dot test %s: %.3e" % (precision, abs(lhs - rhs) / abs(lhs)))
    assert abs(lhs - rhs) < tol * abs(lhs)
    torch.cuda.synchronize()


@pytest.mark.parametrize("nx,nw", [(1024, 4), (4096, 2), (8192, 1)])
def test_transforms_and_fused_passes_are_bitwise_deterministic(cuda, nx, nw):
    """Run-to-run determinism of every pass that uses raw buffer loads and stores or
    direct-to-LDS loads: the plane transforms (both signs, both precisions), the fused
    imaging path (FP64 strip-major scratch and FP32C row-major, column transform over
    empty tiles, fused or split row pass) and the adjoint give the same bits when run
    twice on the same input."""
    import torch
    from lofar_amd import imager
    rng = np.random.default_rng(nx + nw)
    x = rng.standard_normal((nw, nx, nx)) + 1j * rng.standard_normal((nw, nx, nx))
    for dt in (torch.complex128, torch.complex64):
        c = torch.from_numpy(x).to(cuda).to(dt)
        for sign in (1, -1):
            a = imager.fft(c, sign=sign)
            b = imager.fft(c, sign=sign)
            assert torch.equal(a.view(torch.float32 if dt == torch.complex64 else torch.float64),
                               b.view(torch.float32 if dt == torch.complex64 else torch.float64))
        del c
    del x
    d = synth(300000, nx, max(nw, 2), seed=nx)
    nww = max(nw, 2)
    for prec in ("fp64", "fp32c"):
        outs = []
        for _ in range(2):
            r, i = imager.image(*dev_inputs(d, cuda), nww, 7, nx, nx, d["resolution"], d["wmin"],
                                d["wmax"], precision=prec)
            outs.append((H(r), H(i)))
        assert np.array_equal(outs[0][0].view(np.uint64), outs[1][0].view(np.uint64))
        assert np.array_equal(outs[0][1].view(np.uint64), outs[1][1].view(np.uint64))
        img_re = torch.from_numpy(outs[0][0]).to(cuda)
        img_im = torch.from_numpy(outs[0][1]).to(cuda)
        ms = [imager.degrid(img_re, img_im, T(d["uu"], cuda), T(d["vv"], cuda), T(d["ww"], cuda),
                            T(d["weight"], cuda), nww, 7, d["resolution"], d["wmin"], d["wmax"],
                            precision=prec) for _ in range(2)]
        assert torch.equal(ms[0][0], ms[1][0]) and torch.equal(ms[0][1], ms[1][1])
    torch.cuda.synchronize()


@pytest.mark.parametrize("support", [19, 25, 40])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_wide_support_grid_matches_oracle(cuda, support, precision, tol):
    """Supports whose window (2K+1) exceeds the tile kernels' 17 cells take the bucket-free
    path (global atomics, as the reference's OpenMP loop); windows clipped at every edge,
    NaN real parts, F = 2 channels; bad samples are rejected with nothing written."""
    import torch
    from lofar_amd import imager
    from lofar_amd._lib import ECOORD, LofarError
    d = random_case(40 + support, 6000, 160, 96, 3, F=2, edge=True)
    st, want = O.wstack(3, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / 160, 1 / 3, support, 160, 96, freq_per_chan=2)
    got = H(imager.grid(*dev_inputs(d, cuda), 3, support, 160, 96, precision=precision,
                        freq_per_chan=2))
    assert st == O.OK and rel(got, want) < tol
    # accumulate mode adds to the cube
    c = imager.grid(*dev_inputs(d, cuda), 3, support, 160, 96, precision=precision,
                    freq_per_chan=2)
    c2 = imager.grid(*dev_inputs(d, cuda), 3, support, 160, 96, precision=precision,
                     freq_per_chan=2, cube=c.clone(), accumulate=True)
    assert rel(H(c2), 2 * want) < tol
    bad = dict(d)
    bad["ww"] = d["ww"].copy()
    bad["ww"][123] = 1.5
    keep = torch.full_like(c, 5.0)
    with pytest.raises(LofarError) as ei:
        imager.grid(*dev_inputs(bad, cuda), 3, support, 160, 96, precision=precision,
                    freq_per_chan=2, cube=keep)
    assert ei.value.code == ECOORD and bool((keep == 5.0).all())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-6)])
def test_wide_support_image_and_adjoint(cuda, precision, tol):
    """The fused imaging path with support 21 (bucket-free gridding into a row-major cube)
    against the oracle, the degridder against the oracle's interpolation, and the dot test."""
    from lofar_amd import imager
    nx, nw, support = 256, 4, 21
    d = synth(20000, nx, nw, seed=21)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                   support, nx, nx, d["resolution"], d["wmin"], d["wmax"])
    r, i = imager.image(*dev_inputs(d, cuda), nw, support, nx, nx, d["resolution"], d["wmin"],
                        d["wmax"], precision=precision)
    ax = H(r) + 1j * H(i)
    assert rel(ax, want) < tol
    rng = np.random.default_rng(22)
    # y = Ax (normalised) plus noise, so that <y, Ax> is not a near-cancelling sum
    y = ax / np.sqrt(np.mean(np.abs(ax) ** 2)) + (rng.standard_normal((nx, nx)) +
                                                   1j * rng.standard_normal((nx, nx)))
    want_m = O.degrid_image(y, nw, d["uu"], d["vv"], d["ww"], d["weight"], support,
                            d["resolution"], d["wmin"], d["wmax"])
    mr, mi = imager.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), nw, support,
                           d["resolution"], d["wmin"], d["wmax"], precision=precision)
    ahy = H(mr).astype(np.float64) + 1j * H(mi).astype(np.float64)
    assert rel(ahy, want_m) < (1e-10 if precision == "fp64" else 1e-5)
    x = d["vis_real"].astype(np.float64) + 1j * d["vis_img"].astype(np.float64)
    lhs, rhs = np.vdot(y, ax), np.vdot(ahy, x)
    assert abs(lhs - rhs) < (1e-12 if precision == "fp64" else 1e-6) * abs(lhs)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-6)])
def test_more_super_tiles_than_lds_histograms(cuda, precision, tol):
    """9000 w-planes of 64x64 cells make 9000 super-tiles, more than the 8192 the LDS
    level-1 histograms hold: the bucket sort takes its global-histogram path and the
    image and the degridding still match the oracle."""
    from lofar_amd import imager
    nx, nw = 64, 9000
    d = synth(60000, nx, nw, seed=9000)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7,
                   nx, nx, d["resolution"], d["wmin"], d["wmax"])
    r, i = imager.image(*dev_inputs(d, cuda), nw, 7, nx, nx, d["resolution"], d["wmin"],
                        d["wmax"], precision=precision)
    assert rel(H(r) + 1j * H(i), want) < tol
    rng = np.random.default_rng(9001)
    y = rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx))
    want_m = O.degrid_image(y, nw, d["uu"], d["vv"], d["ww"], d["weight"], 7, d["resolution"],
                            d["wmin"], d["wmax"])
    mr, mi = imager.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), T(d["uu"], cuda),
                           T(d["vv"], cuda), T(d["ww"], cuda), T(d["weight"], cuda), nw, 7,
                           d["resolution"], d["wmin"], d["wmax"], precision=precision)
    assert rel(H(mr) + 1j * H(mi), want_m) < (1e-10 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_grid_bright_outlier_keeps_faint_cells_precise(cuda, precision, tol):
    """One visibility 1e6 times brighter than the rest: the fixed-point quantum of its tile
    would follow the outlier (and cost the faint cells ~1e-8 relative in FP64 mode), so such
    a tile is summed in floating point; the cells outside the outlier's window stay within
    the mode's tolerance of the oracle."""
    from lofar_amd import imager
    n, nx, nw = 20000, 256, 2
    d = random_case(50, n, nx, nx, nw, edge=False, nan=0)
    i0 = 777
    d["vis_real"] = d["vis_real"].copy()
    d["vis_real"][i0] = 1.0e6
    got = H(imager.grid(*dev_inputs(d, cuda), nw, 7, nx, nx, precision=precision))
    st, want = O.wstack(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"],
                        1 / nx, 1 / nw, 7, nx, nx)
    assert st == O.OK
    gu, gv, gw = int(d["uu"][i0] * nx), int(d["vv"][i0] * nx), int(d["ww"][i0] * nw)
    faint = np.ones(want.shape, bool)
    faint[gw, max(gv - 3, 0):gv + 4, max(gu - 3, 0):gu + 4] = False
    # the faint cells of the outlier's own tile, and all of them
    t = np.zeros(want.shape, bool)
    t[gw, (gv // 64) * 64:(gv // 64) * 64 + 64, (gu // 64) * 64:(gu // 64) * 64 + 64] = True
    assert rel(got[faint & t], want[faint & t]) < tol
    assert rel(got[faint], want[faint]) < tol
    assert rel(got, want) < tol
