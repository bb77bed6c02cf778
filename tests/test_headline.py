"""Parity of the HIP path at the headline workloads, against the reference CPU path.

The reference path (oracle/refpath.py) is the reference's own C compiled unchanged
(w-stacking.cu's OpenMP gridder, w-stacking.cu:222-307, and phase_correction.cu's OpenMP
loop, phase_correction.cu:145-210) around a per-plane FFTW_BACKWARD/(Nx*Ny) transform
(fourier_transform.c:31-61) done with scipy.fft.ifft2, run on the host's cores on the
same seeded samples the GPU images.  Tolerances, as relative L2 over the complex image:
FP64 mode 1e-10, FP32C mode 1e-6.  Where a reference run at full size is out of reach
(8192^2 x 128 planes: a 137 GB FP64 cube, and the reference's int indexing overflows),
size-independent properties are checked at full size instead: FP32C against FP64 on the
same samples, the adjoint dot test and a monotone chi^2 over major cycles.

Inputs are drawn on the device with the seeded generator of lofar_amd.synth
(backend="torch"), seeds as bench.py uses them.
"""
import math

import numpy as np
import pytest

from oracle import refpath

pytestmark = pytest.mark.gpu

FIELDS = ("uu", "vv", "ww", "vis_real", "vis_img", "weight")


def _obs(n, nx, nw, seed, dev):
    from lofar_amd.synth import make_obs
    return make_obs(n, nx, nw, seed=seed, backend="torch", device=dev)


def _host(d):
    return {k: d[k].cpu().numpy() for k in FIELDS}


def _image(d, nw, support, nx, precision, cube=None):
    from lofar_amd import imager
    r, i = imager.image(*[d[k] for k in FIELDS], nw, support, nx, nx, d["resolution"],
                        d["wmin"], d["wmax"], precision=precision, cube=cube)
    return r.cpu().numpy() + 1j * i.cpu().numpy()


def _free():
    import torch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def _reference(d, nw, support, nx):
    want, t = refpath.reference_image(_host(d), nw, support, nx, nx, d["resolution"], d["wmin"],
                                      d["wmax"])
    print("reference path: %s; %s; gridding %.1f s, fft %.1f s, phase %.1f s"
          % (t["gridder"], t["phase"], t["kernel_s"], t["fftw_s"], t["phase_s"]), flush=True)
    return want


def test_config2_fp64_16M_4096x32_matches_reference(cuda):
    """16M samples, 4096x4096, 32 w-planes, support 7 (the FP64 drop-in workload): the FP64
    mode within 1e-10 of the reference path, and the FP32C mode within 1e-6."""
    n, nx, nw, support = 16 * 1024 * 1024, 4096, 32, 7
    d = _obs(n, nx, nw, 2024, cuda)
    got64 = _image(d, nw, support, nx, "fp64")
    _free()
    got32 = _image(d, nw, support, nx, "fp32c")
    _free()
    want = _reference(d, nw, support, nx)
    e64 = refpath.image_errors(got64, want)
    e32 = refpath.image_errors(got32, want)
    print("config 2 vs reference: fp64 rms %.3e max %.3e; fp32c rms %.3e max %.3e"
          % (e64 + e32), flush=True)
    assert e64[0] <= 1e-10
    assert e32[0] <= 1e-6


def test_config3_fp32c_64M_4096x32_matches_reference(cuda):
    """64M samples, 4096x4096, 32 w-planes, support 7 (the FP32C precision workload): the
    FP32C mode within 1e-6 of the reference path, the FP64 mode within 1e-10."""
    n, nx, nw, support = 64 * 1024 * 1024, 4096, 32, 7
    d = _obs(n, nx, nw, 2025, cuda)
    got32 = _image(d, nw, support, nx, "fp32c")
    _free()
    got64 = _image(d, nw, support, nx, "fp64")
    _free()
    want = _reference(d, nw, support, nx)
    e32 = refpath.image_errors(got32, want)
    e64 = refpath.image_errors(got64, want)
    print("config 3 vs reference: fp32c rms %.3e max %.3e; fp64 rms %.3e max %.3e"
          % (e32 + e64), flush=True)
    assert e32[0] <= 1e-6
    assert e64[0] <= 1e-10


def test_8192x2_fp64_matches_reference(cuda):
    """8192-wide planes in FP64: the 16-element complex-double column transform with the
    second occupancy word (tile rows 64..127 exist only at 8192), the split row transform
    and phase sum, against the reference path; 2M samples on 2 w-planes, support 7."""
    n, nx, nw, support = 2 * 1024 * 1024, 8192, 2, 7
    d = _obs(n, nx, nw, 8192, cuda)
    got64 = _image(d, nw, support, nx, "fp64")
    _free()
    got32 = _image(d, nw, support, nx, "fp32c")
    _free()
    want = _reference(d, nw, support, nx)
    e64 = refpath.image_errors(got64, want)
    e32 = refpath.image_errors(got32, want)
    print("8192 x 2 vs reference: fp64 rms %.3e max %.3e; fp32c rms %.3e max %.3e"
          % (e64 + e32), flush=True)
    assert e64[0] <= 1e-10
    assert e32[0] <= 1e-6


def _sky(nx, ny, dev, g):
    """The differentiation workload's sky: 3 Gaussian sources plus 1e-3 white noise."""
    import torch
    yy, xx = torch.meshgrid(torch.arange(ny, device=dev, dtype=torch.float64),
                            torch.arange(nx, device=dev, dtype=torch.float64), indexing="ij")
    sky = 1e-3 * torch.randn((ny, nx), generator=g, dtype=torch.float64, device=dev)
    for (cy, cx, amp, sig) in ((0.29 * ny, 0.22 * nx, 1.0, 3.0), (0.71 * ny, 0.61 * nx, 0.6, 5.0),
                               (0.17 * ny, 0.81 * nx, 1.5, 2.0)):
        sky += amp * torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sig * sig))
    return sky


def _dot_test(op, x_re, x_im, g):
    """<y, A^H x> = <A y, x> with A^H the imaging path and A the degridder: y = A^H x
    normalised plus unit noise (so that <y, A^H x> is not a zero-mean sum)."""
    import torch
    ar, ai = op.adjoint(x_re, x_im)
    rms = torch.sqrt(torch.mean(ar * ar + ai * ai))
    yr = ar / rms + torch.randn(ar.shape, generator=g, dtype=torch.float64, device=ar.device)
    yi = ai / rms + torch.randn(ar.shape, generator=g, dtype=torch.float64, device=ar.device)
    mr, mi = op.forward(yr, yi)
    xr, xi = x_re.double(), x_im.double()
    lhs_re = float((yr * ar + yi * ai).sum())
    lhs_im = float((yr * ai - yi * ar).sum())
    hr, hi = mr.double(), mi.double()
    rhs_re = float((hr * xr + hi * xi).sum())
    rhs_im = float((hr * xi - hi * xr).sum())
    return math.hypot(lhs_re - rhs_re, lhs_im - rhs_im) / math.hypot(lhs_re, lhs_im)


def test_config4_adjoint_16M_4096x32(cuda):
    """Differentiation workload: degridding a 4096x4096 sky (3 Gaussian sources plus 1e-3
    noise) to 16M samples on 32 w-planes, support 7.  The dot test in both modes (1e-6),
    the chi^2 gradient of the torch.autograd binding against central finite differences on
    64 random pixels (FP64 mode, the loss is quadratic so any step is exact up to
    rounding), the gradient against -2x the reference path's dirty image of the weighted
    residuals, and 5 monotone gradient-descent steps."""
    import torch
    from lofar_amd.autograd import ImagingOperator, chi2, chi2_descent, measure
    n, nx, nw, support = 16 * 1024 * 1024, 4096, 32, 7
    d = _obs(n, nx, nw, 2024, cuda)
    g = torch.Generator(device=cuda)
    g.manual_seed(4)
    for prec in ("fp32c", "fp64"):
        op = ImagingOperator(d["uu"], d["vv"], d["ww"], None, nw, support, nx, nx,
                             d["resolution"], d["wmin"], d["wmax"], precision=prec)
        e = _dot_test(op, d["vis_real"], d["vis_img"], g)
        print("config 4 dot test %s: %.3e" % (prec, e), flush=True)
        assert e <= 1e-6
        del op
        _free()
    op = ImagingOperator(d["uu"], d["vv"], d["ww"], None, nw, support, nx, nx, d["resolution"],
                         d["wmin"], d["wmax"], precision="fp64")
    w = d["weight"]
    xr = _sky(nx, nx, cuda, g).requires_grad_(True)
    xi = torch.zeros_like(xr).requires_grad_(True)
    m_re, m_im = measure(xr, xi, op)
    loss = chi2(m_re, m_im, d["vis_real"], d["vis_img"], w)
    loss.backward()
    rng = np.random.default_rng(44)
    pix = rng.choice(nx * nx, 64, replace=False)
    h = 100.0
    fd, ad = [], []
    with torch.no_grad():
        for k, p in enumerate(pix):
            x = xr if k % 2 == 0 else xi
            gr = xr.grad if k % 2 == 0 else xi.grad
            v, u = divmod(int(p), nx)
            x0 = float(x[v, u])
            x[v, u] = x0 + h
            lp = float(chi2(*op.forward(xr, xi), d["vis_real"], d["vis_img"], w))
            x[v, u] = x0 - h
            lm = float(chi2(*op.forward(xr, xi), d["vis_real"], d["vis_img"], w))
            x[v, u] = x0
            fd.append((lp - lm) / (2 * h))
            ad.append(float(gr[v, u]))
    fd, ad = np.array(fd), np.array(ad)
    err_fd = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print("config 4 central differences on 64 pixels (fp64): rel %.3e" % err_fd, flush=True)
    assert err_fd <= 1e-6
    # the gradient is -2 x the dirty image of the weighted residuals (float32, as the
    # reference's wstack() takes visibilities), dirty image from the reference path
    r_re = (d["vis_real"].double() - m_re.detach().double()).float()
    r_im = (d["vis_img"].double() - m_im.detach().double()).float()
    res = dict(d, vis_real=r_re, vis_img=r_im)
    got = xr.grad.cpu().numpy() + 1j * xi.grad.cpu().numpy()
    del m_re, m_im, r_re, r_im
    _free()
    dirty = _reference(res, nw, support, nx)
    err_g = np.linalg.norm(got + 2 * dirty) / np.linalg.norm(2 * dirty)
    print("config 4 gradient vs -2 x reference dirty image of residuals: rel %.3e" % err_g,
          flush=True)
    assert err_g <= 1e-6
    hist = []
    chi2_descent(op, d["vis_real"], d["vis_img"], w, xr.detach(), xi.detach(), iters=5,
                 history=hist)
    hv = np.array([float(v) for v in hist])
    print("config 4 chi2 over 5 steps: %s" % np.array2string(hv, precision=6), flush=True)
    assert np.all(np.diff(hv) <= 1e-10 * hv[0])


def test_config5_8192x128_s9_256M_properties(cuda):
    """256M samples (16M baselines x 16 channels, 120-168 MHz, uvw scaled by nu/c) streamed
    to the device in 32M-sample chunks, 8192x8192 grid, 128 w-planes, support 9: at full
    size, the FP32C image within 1e-6 of the FP64 mode on the same samples (a reference run
    is out of reach: 137 GB FP64 cube, int indexing overflow in the reference), the adjoint
    dot test within 1e-6, and chi^2 decreasing over 5 major cycles."""
    import torch
    from lofar_amd import imager
    from lofar_amd.autograd import ImagingOperator, lipschitz
    from lofar_amd.stream import major_cycles, make_multichannel
    nx, nw, support = 8192, 128, 9
    obs = make_multichannel(16 * 1024 * 1024, nx, nchan=16, seed=2026, device=cuda)
    s, secs, nbytes = obs.to_device(cuda, chunk_samples=32 * 1024 * 1024)
    print("config 5: %d samples streamed (%.2f GB in %.2f s)" % (obs.num_samples, nbytes / 1e9,
                                                                secs), flush=True)
    res, wmin, wmax = 1.0 / int(obs.span), obs.wmin, obs.wmax
    del obs
    vis = [s[k] for k in FIELDS]
    args = (nw, support, nx, nx, res, wmin, wmax)
    cube = torch.empty((nw, nx, nx), dtype=torch.complex64, device=cuda)
    r32, i32 = imager.image(*vis, *args, precision="fp32c", cube=cube)
    # dot test and major cycles on the FP32C cube
    op = ImagingOperator(s["uu"], s["vv"], s["ww"], None, nw, support, nx, nx, res, wmin, wmax,
                         precision="fp32c", cube=cube)
    g = torch.Generator(device=cuda)
    g.manual_seed(5)
    e_dot = _dot_test(op, s["vis_real"], s["vis_img"], g)
    print("config 5 dot test fp32c: %.3e" % e_dot, flush=True)
    step = 0.5 / lipschitz(op, s["weight"], iters=3)
    x0 = torch.zeros((nx, nx), dtype=torch.float64, device=cuda)
    hist = []
    major_cycles(op, s["vis_real"], s["vis_img"], s["weight"], x0, x0.clone(), cycles=5,
                 step=step, history=hist)
    hv = np.array([float(v) for v in hist])
    print("config 5 chi2 over 5 major cycles: %s" % np.array2string(hv, precision=6), flush=True)
    del op, cube, x0
    _free()
    c64 = torch.empty((nw, nx, nx), dtype=torch.complex128, device=cuda)
    r64, i64 = imager.image(*vis, *args, precision="fp64", cube=c64)
    del c64
    _free()
    e32 = refpath.image_errors(r32.cpu().numpy() + 1j * i32.cpu().numpy(),
                               r64.cpu().numpy() + 1j * i64.cpu().numpy())
    print("config 5 fp32c vs fp64: rms %.3e max %.3e" % e32, flush=True)
    assert e32[0] <= 1e-6
    assert e_dot <= 1e-6
    assert np.all(np.diff(hv) <= 1e-6 * hv[0]) and hv[-1] < hv[0]
