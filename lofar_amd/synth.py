"""Seeded synthetic observations for the imaging hot path.

Layout and units follow the reference's inputs (init.c:390-494, w-stacking_omp.c:85-90):
uu, vv, ww are float64 normalised to [0, 1); vis_real/vis_img float32 [N*F*P];
weights float32 [N*P].  The sky and sampling model:

* (u, v) uniform in a disk of radius 0.45 around (0.5, 0.5) in normalised units;
  the physical span is S_uv = 2*Nx wavelengths, u_lambda = (u_n - 0.5)*S_uv, and the
  metadata uvmin=0, uvmax=S_uv give resolution = 1/S_uv (gridding.c:160).
* w_lambda ~ N(0, 0.1*max|uv_lambda|) clipped to +-3 sigma; wmin/wmax = -+3 sigma and
  w_n = (w_lambda - wmin)/(wmax - wmin), clamped below 1.
* V = sum_s S*exp(-2 pi^2 sigma^2 (u^2+v^2)) * exp(-2 pi i (u l + v m + w (n-1)))
  over Gaussian sources with |l|,|m| < 0.2, S ~ U(0.5, 2), sigma ~ U(1e-3, 5e-3) rad,
  plus complex N(0, noise^2); weights 1.

``backend="numpy"`` draws with numpy.random.default_rng(seed) (test sizes);
``backend="torch"`` draws the same model with a seeded torch generator on the
given device (full-size benchmark inputs).
"""
from __future__ import annotations

import math

import numpy as np

ONE_MINUS = 1.0 - 2.0 ** -52


def _sources(rng_uniform, n_sources):
    srcs = []
    for _ in range(n_sources):
        l, m = (rng_uniform(-0.2, 0.2), rng_uniform(-0.2, 0.2))
        srcs.append(dict(l=l, m=m, flux=rng_uniform(0.5, 2.0), sigma=rng_uniform(1e-3, 5e-3)))
    return srcs


def make_obs(n, nx, nw, seed=0, backend="numpy", device=None, n_sources=2, noise=0.01,
             radius=0.45):
    """Returns a dict with uu, vv, ww, vis_real, vis_img, weight and the metadata
    (wmin, wmax, uvmin, uvmax, resolution, span, sources)."""
    span = 2.0 * nx
    if backend == "numpy":
        rng = np.random.default_rng(seed)
        srcs = _sources(lambda a, b: float(rng.uniform(a, b)), n_sources)
        r = radius * np.sqrt(rng.random(n))
        th = 2 * np.pi * rng.random(n)
        un = 0.5 + r * np.cos(th)
        vn = 0.5 + r * np.sin(th)
        ul = (un - 0.5) * span
        vl = (vn - 0.5) * span
        uvmax_l = float(np.sqrt(ul * ul + vl * vl).max()) if n else radius * span
        sig_w = 0.1 * uvmax_l
        wl = np.clip(rng.normal(0.0, sig_w, n), -3 * sig_w, 3 * sig_w)
        wmin, wmax = -3 * sig_w, 3 * sig_w
        wn = np.clip((wl - wmin) / (wmax - wmin), 0.0, ONE_MINUS)
        vis = np.zeros(n, np.complex128)
        for s in srcs:
            nn = math.sqrt(1 - s["l"] ** 2 - s["m"] ** 2)
            amp = s["flux"] * np.exp(-2 * np.pi ** 2 * s["sigma"] ** 2 * (ul * ul + vl * vl))
            vis += amp * np.exp(-2j * np.pi * (ul * s["l"] + vl * s["m"] + wl * (nn - 1)))
        vis += noise * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
        out = dict(uu=un, vv=vn, ww=wn, vis_real=vis.real.astype(np.float32),
                   vis_img=vis.imag.astype(np.float32), weight=np.ones(n, np.float32))
    elif backend == "torch":
        import torch
        dev = torch.device(device or "cuda")
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        hg = np.random.default_rng(seed)
        srcs = _sources(lambda a, b: float(hg.uniform(a, b)), n_sources)
        f64 = dict(dtype=torch.float64, device=dev)
        r = radius * torch.sqrt(torch.rand(n, generator=g, **f64))
        th = 2 * math.pi * torch.rand(n, generator=g, **f64)
        un = 0.5 + r * torch.cos(th)
        vn = 0.5 + r * torch.sin(th)
        del r, th
        ul = (un - 0.5) * span
        vl = (vn - 0.5) * span
        uvmax_l = float(torch.sqrt(ul * ul + vl * vl).max()) if n else radius * span
        sig_w = 0.1 * uvmax_l
        wl = torch.clamp(torch.randn(n, generator=g, **f64) * sig_w, -3 * sig_w, 3 * sig_w)
        wmin, wmax = -3 * sig_w, 3 * sig_w
        wn = torch.clamp((wl - wmin) / (wmax - wmin), 0.0, ONE_MINUS)
        vr = torch.zeros(n, **f64)
        vi = torch.zeros(n, **f64)
        q2 = ul * ul + vl * vl
        for s in srcs:
            nn = math.sqrt(1 - s["l"] ** 2 - s["m"] ** 2)
            amp = s["flux"] * torch.exp(-2 * math.pi ** 2 * s["sigma"] ** 2 * q2)
            ph = -2 * math.pi * (ul * s["l"] + vl * s["m"] + wl * (nn - 1))
            vr += amp * torch.cos(ph)
            vi += amp * torch.sin(ph)
            del amp, ph
        vr += noise * torch.randn(n, generator=g, **f64)
        vi += noise * torch.randn(n, generator=g, **f64)
        out = dict(uu=un, vv=vn, ww=wn, vis_real=vr.float(), vis_img=vi.float(),
                   weight=torch.ones(n, dtype=torch.float32, device=dev))
        del vr, vi, q2, ul, vl, wl
    else:
        raise ValueError(backend)
    out.update(wmin=wmin, wmax=wmax, uvmin=0.0, uvmax=span, span=span,
               resolution=1.0 / int(span), sources=srcs, nx=nx, nw=nw)
    return out
