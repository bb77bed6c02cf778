"""torch.autograd harness of the imaging operator (lofar_amd.autograd) on the device.

Checks: torch.autograd.gradcheck on a small problem, central finite differences of the
weighted chi^2 on 64 random pixels, the chi^2 gradient against the oracle's dirty image of
the weighted residuals, and monotone gradient descent.  The loss is quadratic in the image,
so a central difference equals the directional derivative for any step h up to rounding.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _obs(n, nx, nw, seed, dev):
    import torch
    from lofar_amd.synth import make_obs
    d = make_obs(n, nx, nw, seed=seed)
    t = {k: torch.from_numpy(np.ascontiguousarray(d[k])).to(dev)
         for k in ("uu", "vv", "ww", "vis_real", "vis_img", "weight")}
    return d, t


def _op(t, d, nw, nx, precision, weight=None):
    from lofar_amd.autograd import ImagingOperator
    return ImagingOperator(t["uu"], t["vv"], t["ww"], weight, nw, 7, nx, nx, d["resolution"],
                           d["wmin"], d["wmax"], precision=precision)


def test_gradcheck_fp64(cuda):
    import torch
    from lofar_amd.autograd import measure
    d, t = _obs(120, 32, 3, 31, cuda)
    op = _op(t, d, 3, 32, "fp64", weight=t["weight"] * 0.5 + 0.25)
    g = torch.Generator(device=cuda)
    g.manual_seed(32)
    xr = torch.randn((32, 32), generator=g, dtype=torch.float64, device=cuda, requires_grad=True)
    xi = torch.randn((32, 32), generator=g, dtype=torch.float64, device=cuda, requires_grad=True)
    # the backward pass sums in LDS with FP64 atomics: reruns agree to rounding, not bitwise
    assert torch.autograd.gradcheck(lambda a, b: measure(a, b, op), (xr, xi), eps=1e-3,
                                    atol=1e-9, rtol=1e-7, nondet_tol=1e-12)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("fp32c", 1e-4)])
def test_chi2_gradient_matches_central_differences(cuda, precision, tol):
    import torch
    from lofar_amd.autograd import chi2, measure
    nx, nw = 256, 8
    d, t = _obs(40000, nx, nw, 33, cuda)
    op = _op(t, d, nw, nx, precision)
    g = torch.Generator(device=cuda)
    g.manual_seed(34)
    xr = (1e-3 * torch.randn((nx, nx), generator=g, dtype=torch.float64, device=cuda))
    xi = (1e-3 * torch.randn((nx, nx), generator=g, dtype=torch.float64, device=cuda))
    xr.requires_grad_(True)
    xi.requires_grad_(True)
    w = t["weight"]
    loss = chi2(*measure(xr, xi, op), t["vis_real"], t["vis_img"], w)
    loss.backward()
    rng = np.random.default_rng(35)
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
            lp = float(chi2(*op.forward(xr, xi), t["vis_real"], t["vis_img"], w))
            x[v, u] = x0 - h
            lm = float(chi2(*op.forward(xr, xi), t["vis_real"], t["vis_img"], w))
            x[v, u] = x0
            fd.append((lp - lm) / (2 * h))
            ad.append(float(gr[v, u]))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print("finite differences %s: rel %.2e" % (precision, err))
    assert err < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("fp32c", 1e-5)])
def test_chi2_gradient_is_dirty_image_of_residuals(cuda, precision, tol):
    """d chi^2/dx = -2 * dirty image of the weighted residuals, the dirty image taken from the
    oracle path (wstack -> FFTW_BACKWARD/(NxNy) -> phase_correction) on float32 residuals,
    as the reference's gridder takes them."""
    import torch
    from lofar_amd.autograd import chi2, measure
    nx, nw = 256, 4
    d, t = _obs(30000, nx, nw, 36, cuda)
    op = _op(t, d, nw, nx, precision)
    g = torch.Generator(device=cuda)
    g.manual_seed(37)
    xr = (1e-3 * torch.randn((nx, nx), generator=g, dtype=torch.float64, device=cuda))
    xi = torch.zeros_like(xr)
    xr.requires_grad_(True)
    xi.requires_grad_(True)
    m_re, m_im = measure(xr, xi, op)
    loss = chi2(m_re, m_im, t["vis_real"], t["vis_img"], t["weight"])
    loss.backward()
    r_re = (t["vis_real"].double() - m_re.detach().double()).cpu().numpy().astype(np.float32)
    r_im = (t["vis_img"].double() - m_im.detach().double()).cpu().numpy().astype(np.float32)
    dirty = O.image(nw, d["uu"], d["vv"], d["ww"], r_re, r_im, d["weight"], 7, nx, nx,
                    d["resolution"], d["wmin"], d["wmax"])
    got = xr.grad.cpu().numpy() + 1j * xi.grad.cpu().numpy()
    err = np.linalg.norm(got + 2 * dirty) / np.linalg.norm(2 * dirty)
    print("gradient vs dirty image %s: rel %.2e" % (precision, err))
    assert err < tol


@pytest.mark.parametrize("precision,noise", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_chi2_descent_decreases_monotonically(cuda, precision, noise):
    import torch
    from lofar_amd.autograd import chi2_descent
    nx, nw = 256, 8
    d, t = _obs(50000, nx, nw, 38, cuda)
    op = _op(t, d, nw, nx, precision)
    x0 = torch.zeros((nx, nx), dtype=torch.float64, device=cuda)
    hist = []
    chi2_descent(op, t["vis_real"], t["vis_img"], t["weight"], x0, x0.clone(), iters=20,
                 history=hist)
    h = np.array([float(v) for v in hist])
    print("chi2 %s: %.6e -> %.6e" % (precision, h[0], h[-1]))
    assert np.all(np.diff(h) <= noise * h[0])
    assert h[-1] < 0.9 * h[0]
