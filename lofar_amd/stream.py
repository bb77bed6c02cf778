"""Multi-channel observations streamed to the device, and the major-cycle loop.

The reference grids every frequency channel of a sample at one (u, v, w)
(w-stacking_omp.c:116-133).  Here each channel is its own sample: baseline coordinates in
metres are scaled by nu/c to wavelengths per channel and normalised to the reference's
[0, 1) units (u_n = u_lambda/span + 1/2 with span a power of two, w_n = (w_lambda - wmin)
times 1/(wmax - wmin), clamped below 1), then gridded
with F = P = 1.

``ChannelStream`` keeps an observation in pinned host memory (baseline uvw in metres and
per-channel visibilities) and moves it to resident device arrays in chunks of baselines: the
copies run on a side stream while the compute stream expands the previous chunk into
per-channel samples, so transfer and expansion overlap.  ``major_cycles`` then runs the
degrid -> residual -> grid loop (gradient steps on the weighted chi^2 through
lofar_amd.autograd) on the resident samples.
"""
from __future__ import annotations

import math
import time

import numpy as np

C_LIGHT = 299792458.0
ONE_MINUS = 1.0 - 2.0 ** -52


def _torch():
    import torch
    return torch


class ChannelStream:
    """An observation of ``n_base`` baseline samples x ``len(freqs)`` channels.

    base_uvw: float64 [n_base, 3] (metres); vis_re, vis_im: float32 [n_base, nchan];
    freqs: channel frequencies (Hz); span: the uv extent in wavelengths that maps to the
    grid (normalised u = u_lambda/span + 1/2); wmin, wmax: the w range (wavelengths) the
    w-planes cover.  Host tensors are pinned so that the copies are asynchronous."""

    def __init__(self, base_uvw, vis_re, vis_im, freqs, span, wmin, wmax):
        torch = _torch()
        self.base_uvw = base_uvw
        self.vis_re = vis_re
        self.vis_im = vis_im
        self.freqs = torch.as_tensor(np.asarray(freqs, np.float64))
        self.span, self.wmin, self.wmax = float(span), float(wmin), float(wmax)
        self.n_base, self.nchan = vis_re.shape

    @property
    def num_samples(self):
        return self.n_base * self.nchan

    def to_device(self, device, chunk_samples=32 * 1024 * 1024):
        """Streams the observation to device-resident sample arrays (uu, vv, ww float64,
        vis_real, vis_img, weight float32; sample index = baseline*nchan + channel) in
        chunks of about ``chunk_samples`` samples.  Returns (arrays, seconds, host bytes)."""
        torch = _torch()
        n, nc = self.n_base, self.nchan
        out = dict(uu=torch.empty(n * nc, dtype=torch.float64, device=device),
                   vv=torch.empty(n * nc, dtype=torch.float64, device=device),
                   ww=torch.empty(n * nc, dtype=torch.float64, device=device),
                   vis_real=torch.empty(n * nc, dtype=torch.float32, device=device),
                   vis_img=torch.empty(n * nc, dtype=torch.float32, device=device),
                   weight=torch.ones(n * nc, dtype=torch.float32, device=device))
        scale = (self.freqs.to(device) / C_LIGHT).view(1, nc)  # metres -> wavelengths
        rows = max(1, chunk_samples // nc)
        copy = torch.cuda.Stream(device=device)
        comp = torch.cuda.current_stream(device)
        staged = []
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for b0 in range(0, n, rows):
            b1 = min(n, b0 + rows)
            with torch.cuda.stream(copy):
                uvw = self.base_uvw[b0:b1].to(device, non_blocking=True)
                vr = self.vis_re[b0:b1].to(device, non_blocking=True)
                vi = self.vis_im[b0:b1].to(device, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(copy)
            staged.append((b0, b1, uvw, vr, vi, ev))
            if len(staged) == 2:
                self._expand(out, scale, comp, *staged.pop(0))
        while staged:
            self._expand(out, scale, comp, *staged.pop(0))
        torch.cuda.synchronize(device)
        secs = time.perf_counter() - t0
        nbytes = n * (3 * 8 + 2 * 4 * nc)
        return out, secs, nbytes

    def _expand(self, out, scale, comp, b0, b1, uvw, vr, vi, ev):
        torch = _torch()
        comp.wait_event(ev)
        for t in (uvw, vr, vi):
            t.record_stream(comp)
        nc = self.nchan
        s0, s1 = b0 * nc, b1 * nc
        ul = uvw[:, 0:1] * scale
        vl = uvw[:, 1:2] * scale
        wl = uvw[:, 2:3] * scale
        out["uu"][s0:s1] = (ul / self.span + 0.5).reshape(-1)
        out["vv"][s0:s1] = (vl / self.span + 0.5).reshape(-1)
        wn = (wl - self.wmin) * (1.0 / (self.wmax - self.wmin))
        out["ww"][s0:s1] = torch.clamp(wn, 0.0, ONE_MINUS).reshape(-1)
        out["vis_real"][s0:s1] = vr.reshape(-1)
        out["vis_img"][s0:s1] = vi.reshape(-1)


def make_multichannel(n_base, nx, nchan=16, f_lo=120e6, f_hi=168e6, seed=0, n_sources=2,
                      noise=0.01, radius=0.45, device=None, gen_chunk=4 * 1024 * 1024):
    """Seeded synthetic multi-channel observation in pinned host memory.

    Baselines uniform in a disk whose longest baseline at f_hi spans ``radius`` of the grid's
    uv extent (span = 2*nx wavelengths); w_metres ~ N(0, 0.1*max|uv|) clipped to 3 sigma;
    the sky and noise model of lofar_amd.synth, evaluated per channel in wavelengths.  The
    visibilities are computed on the device in chunks and copied to the host."""
    torch = _torch()
    from .synth import _sources
    dev = torch.device(device or "cuda")
    span = 2.0 * nx
    lam_min = C_LIGHT / f_hi
    r_m = radius * span * lam_min  # metres: radius*span wavelengths at the top channel
    freqs = np.linspace(f_lo, f_hi, nchan)
    hg = np.random.default_rng(seed)
    srcs = _sources(lambda a, b: float(hg.uniform(a, b)), n_sources)
    sig_w = 0.1 * r_m
    wmax = 3 * sig_w / lam_min
    wmin = -wmax
    base = torch.empty((n_base, 3), dtype=torch.float64, pin_memory=True)
    vre = torch.empty((n_base, nchan), dtype=torch.float32, pin_memory=True)
    vim = torch.empty((n_base, nchan), dtype=torch.float32, pin_memory=True)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    f64 = dict(dtype=torch.float64, device=dev)
    scale = torch.as_tensor(freqs / C_LIGHT, **f64).view(1, nchan)
    for b0 in range(0, n_base, gen_chunk):
        m = min(n_base, b0 + gen_chunk) - b0
        r = r_m * torch.sqrt(torch.rand(m, generator=g, **f64))
        th = 2 * math.pi * torch.rand(m, generator=g, **f64)
        um, vm = r * torch.cos(th), r * torch.sin(th)
        wm = torch.clamp(torch.randn(m, generator=g, **f64) * sig_w, -3 * sig_w, 3 * sig_w)
        base[b0:b0 + m] = torch.stack([um, vm, wm], 1).cpu()
        ul, vl, wl = um.view(-1, 1) * scale, vm.view(-1, 1) * scale, wm.view(-1, 1) * scale
        q2 = ul * ul + vl * vl
        vr = noise * torch.randn((m, nchan), generator=g, **f64)
        vi = noise * torch.randn((m, nchan), generator=g, **f64)
        for s in srcs:
            nn = math.sqrt(1 - s["l"] ** 2 - s["m"] ** 2)
            amp = s["flux"] * torch.exp(-2 * math.pi ** 2 * s["sigma"] ** 2 * q2)
            ph = -2 * math.pi * (ul * s["l"] + vl * s["m"] + wl * (nn - 1))
            vr += amp * torch.cos(ph)
            vi += amp * torch.sin(ph)
        vre[b0:b0 + m] = vr.float().cpu()
        vim[b0:b0 + m] = vi.float().cpu()
    return ChannelStream(base, vre, vim, freqs, span, wmin, wmax)


def major_cycles(op, vis_re, vis_im, weight, x_re, x_im, cycles=5, step=None, history=None):
    """``cycles`` major cycles, each degrid (model visibilities) -> residual -> grid (dirty
    image of the weighted residuals) -> image update, as gradient steps on the weighted
    chi^2 (lofar_amd.autograd.chi2_descent)."""
    from .autograd import chi2_descent
    return chi2_descent(op, vis_re, vis_im, weight, x_re, x_im, iters=cycles, step=step,
                        history=history)
