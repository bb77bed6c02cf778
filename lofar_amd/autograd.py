"""torch.autograd binding of the imaging operator and its adjoint.

The measurement operator A maps a sky image x (complex [ny, nx], held as two float64
tensors) to model visibilities at the samples (u, v, w):

    A = W . G^T . F^H . P^H

P^H is the conjugate w-phase screen per plane times dwnorm*sqrt(1-l^2-m^2) (the adjoint of
phase_correction(), phase_correction.cu:152-199), F^H the forward transform scaled by
1/(Nx*Ny) (the adjoint of fftw_data()'s FFTW_BACKWARD/(Nx*Ny), fourier_transform.c:25-64),
G^T the Gaussian interpolation at each sample with the gridder's window and weights
(the adjoint of wstack(), w-stacking_omp.c:93-141) and W the per-sample weight.  Its
adjoint A^H = P . F . G . W is the imaging path itself: the dirty image of weighted
visibilities.  Forward runs lofar_degrid_dev, backward runs lofar_image_dev, both on the
device (liblofar_hip); there is no other code path.

``measure(img_re, img_im, op)`` is differentiable in both image tensors; the gradient of a
real loss L is (Re, Im) of A^H g with g = dL/dm_re + i dL/dm_im.  For the weighted chi^2
sum_i w_i |V_i - (A0 x)_i|^2 with the unweighted operator A0, the gradient is -2 times the
dirty image of the weighted residuals, and ``chi2_descent`` runs gradient descent on it.

The imaging path takes float32 visibilities, as wstack() does (w-stacking.h:12-29).  In FP64
mode the backward pass therefore splits the float64 cotangent into a float32 high part and a
float32 remainder and images both (the operator is linear), so the adjoint keeps ~2^-48
relative accuracy; the per-sample weight is applied in float64 before the split and the
imaging calls use unit weights.
"""
from __future__ import annotations

from . import imager
from ._lib import FP64

__all__ = ["ImagingOperator", "measure", "chi2", "chi2_descent"]


def _torch():
    import torch
    return torch


class ImagingOperator:
    """A = W G^T F^H P^H for fixed samples and geometry, with one reused device cube.

    uu, vv, ww: float64 CUDA tensors [N] normalised as the reference's inputs;
    weight: float32 CUDA tensor [N] or None (unit weights); cube: optional scratch
    [nw, ny, nx] of the precision's complex dtype.  With ``plan`` (the default) the
    samples are sorted once, at the first application (``imager.Plan``), and every
    later forward and adjoint reuses that sort: the same results without re-sorting
    or host synchronisation; the operator then assumes uu, vv, ww and weight do not
    change."""

    def __init__(self, uu, vv, ww, weight, num_w_planes, w_support, grid_size_x, grid_size_y,
                 resolution, wmin, wmax, precision="fp64", cube=None, plan=True):
        torch = _torch()
        self.uu, self.vv, self.ww = uu, vv, ww
        n = uu.numel()
        self.weight = weight if weight is not None else torch.ones(n, dtype=torch.float32,
                                                                   device=uu.device)
        self.unit = torch.ones(n, dtype=torch.float32, device=uu.device)
        self.nw, self.support = num_w_planes, w_support
        self.nx, self.ny = grid_size_x, grid_size_y
        self.resolution, self.wmin, self.wmax = resolution, wmin, wmax
        self.precision = imager._prec(precision)
        if cube is None:
            cube = torch.empty((num_w_planes, grid_size_y, grid_size_x),
                               dtype=imager.cube_dtype(self.precision), device=uu.device)
        self.cube = cube  # scratch shared by forward and adjoint
        self._use_plan = plan
        self._plan = None

    @property
    def num_samples(self):
        return self.uu.numel()

    def plan(self):
        """The samples' sort (created at the first call), or None without plans."""
        if self._use_plan and self._plan is None:
            self._plan = imager.Plan(self.uu, self.vv, self.ww, self.weight, self.nw,
                                     self.support, self.nx, self.ny, precision=self.precision)
        return self._plan

    def forward(self, img_re, img_im):
        """Model visibilities (float64 in FP64 mode, float32 in FP32C mode)."""
        pl = self.plan()
        if pl is not None:
            return pl.degrid(img_re.contiguous(), img_im.contiguous(), self.resolution,
                             self.wmin, self.wmax, cube=self.cube)
        return imager.degrid(img_re.contiguous(), img_im.contiguous(), self.uu, self.vv, self.ww,
                             self.weight, self.nw, self.support, self.resolution, self.wmin,
                             self.wmax, precision=self.precision, cube=self.cube)

    def _image(self, vr, vi, wt=None, **kw):
        wt = self.unit if wt is None else wt
        pl = self.plan()
        if pl is not None:
            return pl.image(vr, vi, wt, self.resolution, self.wmin, self.wmax,
                            cube=self.cube, **kw)
        return imager.image(self.uu, self.vv, self.ww, vr, vi, wt, self.nw, self.support,
                            self.nx, self.ny, self.resolution, self.wmin, self.wmax,
                            precision=self.precision, cube=self.cube, **kw)

    def adjoint(self, vis_re, vis_im):
        """Dirty image (float64 re, im) of the weighted visibilities: A^H v."""
        torch = _torch()
        if (self.precision != FP64 and vis_re.dtype == torch.float32 and
                vis_im.dtype == torch.float32 and self.weight.dtype == torch.float32):
            # float32 cotangents: the library's float product weight*vis is the correctly
            # rounded exact product, the same value as rounding the double product below
            return self._image(vis_re.contiguous(), vis_im.contiguous(),
                               self.weight.contiguous())
        w = self.weight.to(torch.float64)
        gr = vis_re.to(torch.float64) * w
        gi = vis_im.to(torch.float64) * w
        hr, hi = gr.to(torch.float32), gi.to(torch.float32)
        re, im = self._image(hr.contiguous(), hi.contiguous())
        if self.precision == FP64:
            lr = (gr - hr.to(torch.float64)).to(torch.float32).contiguous()
            li = (gi - hi.to(torch.float64)).to(torch.float32).contiguous()
            self._image(lr, li, image_real=re, image_imag=im, accumulate=True)
        return re, im


def _function():
    torch = _torch()

    class _Measure(torch.autograd.Function):
        @staticmethod
        def forward(ctx, img_re, img_im, op):
            ctx.op = op
            return op.forward(img_re, img_im)

        @staticmethod
        def backward(ctx, g_re, g_im):
            op = ctx.op
            if g_re is None and g_im is None:
                return None, None, None
            n = op.num_samples
            z = None
            if g_re is None or g_im is None:
                like = g_re if g_re is not None else g_im
                z = torch.zeros(n, dtype=like.dtype, device=like.device)
            ir, ii = op.adjoint(g_re if g_re is not None else z, g_im if g_im is not None else z)
            return ir, ii, None

    return _Measure


_MEASURE = None


def measure(img_re, img_im, op):
    """Differentiable model visibilities (m_re, m_im) = A (img_re + i img_im)."""
    global _MEASURE
    if _MEASURE is None:
        _MEASURE = _function()
    return _MEASURE.apply(img_re, img_im, op)


def _chi2_function():
    torch = _torch()

    class _Chi2(torch.autograd.Function):
        """sum_i w_i |V_i - m_i|^2 in float64 and its gradient -2 w (V - m) in one pass each:
        the gradient has the bits of the elementwise autograd graph's (scaling by -2 is
        exact), in a third of its kernels."""

        @staticmethod
        def forward(ctx, m_re, m_im, vis_re, vis_im, weight):
            f64 = torch.float64
            dr = vis_re.to(f64) - m_re
            di = vis_im.to(f64) - m_im
            w = weight.to(f64)
            t = dr * dr
            t.addcmul_(di, di)
            ctx.save_for_backward(dr, di, w)
            return torch.dot(w, t)

        @staticmethod
        def backward(ctx, g):
            dr, di, w = ctx.saved_tensors
            ws = w * (-2.0 * g)
            return ws * dr, ws * di, None, None, None

    return _Chi2


_CHI2 = None


def chi2(m_re, m_im, vis_re, vis_im, weight):
    """Weighted chi^2 sum_i w_i |V_i - m_i|^2, accumulated in float64 (differentiable in
    m_re, m_im)."""
    global _CHI2
    if _CHI2 is None:
        _CHI2 = _chi2_function()
    return _CHI2.apply(m_re, m_im, vis_re, vis_im, weight)


def lipschitz(op, weight, iters=8, seed=0):
    """Power-iteration estimate of the largest eigenvalue of A^H W A (the chi^2 Hessian / 2)."""
    torch = _torch()
    g = torch.Generator(device=op.uu.device)
    g.manual_seed(seed)
    xr = torch.randn((op.ny, op.nx), generator=g, dtype=torch.float64, device=op.uu.device)
    xi = torch.zeros_like(xr)
    lam = 0.0
    w = weight.to(torch.float64)
    for _ in range(iters):
        nrm = torch.sqrt((xr * xr + xi * xi).sum())
        xr, xi = xr / nrm, xi / nrm
        mr, mi = op.forward(xr, xi)
        yr, yi = op.adjoint(mr.to(torch.float64) * w, mi.to(torch.float64) * w)
        lam = float((xr * yr + xi * yi).sum())
        xr, xi = yr, yi
    return lam


def chi2_descent(op, vis_re, vis_im, weight, x_re, x_im, iters=20, step=None, history=None):
    """Gradient descent on the weighted chi^2 of the sky image (x_re, x_im), through
    ``measure``'s autograd; step defaults to 1/(2 L) with L from ``lipschitz`` (monotone
    decrease for this quadratic).  Returns the final image and appends chi^2 per iteration
    (before each step) to ``history``."""
    torch = _torch()
    if step is None:
        step = 0.5 / lipschitz(op, weight)
    x_re = x_re.detach().clone().requires_grad_(True)
    x_im = x_im.detach().clone().requires_grad_(True)
    # float64 copies once, not per iteration
    vis_re, vis_im = vis_re.to(torch.float64), vis_im.to(torch.float64)
    weight = weight.to(torch.float64)
    for _ in range(iters):
        m_re, m_im = measure(x_re, x_im, op)
        loss = chi2(m_re, m_im, vis_re, vis_im, weight)
        if x_re.grad is not None:
            x_re.grad = None
            x_im.grad = None
        loss.backward()
        if history is not None:
            history.append(loss.detach())
        with torch.no_grad():
            x_re -= step * x_re.grad
            x_im -= step * x_im.grad
    return x_re.detach(), x_im.detach()
