"""Host-side mirror of the reference's hot-path interface, on liblofar_hip.

Two layers:

* ``wstack``, ``phase_correction``, ``fftw_data``: the reference's C entry
  points with the same argument names, meaning and accumulate (``+=``)
  semantics (w-stacking.h:12-51, fourier_transform.c:7-74), taking NumPy host
  arrays.  They call the C drop-ins and raise ``LofarError`` where the C
  functions set ``lofar_last_error()``.
* ``grid``, ``fft``, ``phase``, ``image``, ``phase_adjoint``, ``degrid_cube``,
  ``degrid``: the device-resident entry points on torch CUDA tensors, enqueued on
  torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FP32C, FP64, LofarError, check, lib

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _host(a, dtype, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError("%s must be a C-contiguous numpy %s array" % (name, np.dtype(dtype).name))
    return a


def _after_dropin(where):
    code = lib().lofar_last_error()
    if code != 0:
        raise LofarError(where, code)


# ---------------------------------------------------------------------------
# the reference's C interface (host arrays)

def wstack(num_w_planes, num_points, freq_per_chan, polarizations, uu, vv, ww, vis_real, vis_img,
           weight, dx, dw, w_support, grid_size_x, grid_size_y, grid, num_threads=1):
    """wstack() of w-stacking.h:12-29: grids into ``grid`` (float64, 2*Nw*Ny*Nx), accumulating."""
    uu = _host(uu, np.float64, "uu")
    vv = _host(vv, np.float64, "vv")
    ww = _host(ww, np.float64, "ww")
    vr = _host(vis_real, np.float32, "vis_real")
    vi = _host(vis_img, np.float32, "vis_img")
    wt = _host(weight, np.float32, "weight")
    g = _host(grid, np.float64, "grid")
    nfp = num_points * freq_per_chan * polarizations
    if len(uu) < num_points or len(vv) < num_points or len(ww) < num_points:
        raise ValueError("uu/vv/ww shorter than num_points")
    if len(vr) < nfp or len(vi) < nfp or len(wt) < num_points * polarizations:
        raise ValueError("visibility or weight arrays too short")
    if g.size < 2 * num_w_planes * grid_size_x * grid_size_y:
        raise ValueError("grid too small")
    lib().wstack(num_w_planes, num_points, freq_per_chan, polarizations,
                 uu.ctypes.data_as(_dp), vv.ctypes.data_as(_dp), ww.ctypes.data_as(_dp),
                 vr.ctypes.data_as(_fp), vi.ctypes.data_as(_fp), wt.ctypes.data_as(_fp),
                 dx, dw, w_support, grid_size_x, grid_size_y, g.ctypes.data_as(_dp), num_threads)
    _after_dropin("wstack")


def phase_correction(gridss, image_real, image_imag, xaxis, yaxis, num_w_planes, xaxistot,
                     yaxistot, resolution, wmin, wmax, num_threads=1):
    """phase_correction() of w-stacking.h:39-51: accumulates into the two image arrays."""
    g = _host(gridss, np.float64, "gridss")
    re = _host(image_real, np.float64, "image_real")
    im = _host(image_imag, np.float64, "image_imag")
    if g.size < 2 * num_w_planes * xaxis * yaxis or re.size < xaxis * yaxis or im.size < xaxis * yaxis:
        raise ValueError("array too small")
    lib().phase_correction(g.ctypes.data_as(_dp), re.ctypes.data_as(_dp), im.ctypes.data_as(_dp),
                           xaxis, yaxis, num_w_planes, xaxistot, yaxistot, resolution, wmin, wmax,
                           num_threads)
    _after_dropin("phase_correction")


def fftw_data(grid, gridss, grid_size_x, grid_size_y, num_w_planes):
    """The plane transform of fftw_data() (fourier_transform.c:7-74): gridss = iFFT-backward/(NxNy)."""
    g = _host(grid, np.float64, "grid")
    o = _host(gridss, np.float64, "gridss")
    n = 2 * num_w_planes * grid_size_x * grid_size_y
    if g.size < n or o.size < n:
        raise ValueError("array too small")
    check("lofar_fft_planes", lib().lofar_fft_planes(g.ctypes.data_as(_dp), o.ctypes.data_as(_dp),
                                                     grid_size_x, grid_size_y, num_w_planes))


def resolution_from_meta(uvmin, uvmax):
    """gridding.c:160: 1.0/MAX(abs(uvmin),abs(uvmax)) with C's integer abs()."""
    a, b = abs(int(uvmin)), abs(int(uvmax))
    m = max(a, b)
    return float("inf") if m == 0 else 1.0 / m


# ---------------------------------------------------------------------------
# device-resident entry points (torch tensors on the GPU)

def _torch():
    import torch
    return torch


def _ptr(t, dtype, name):
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("%s must be a CUDA tensor" % name)
    if t.dtype != dtype or not t.is_contiguous():
        raise TypeError("%s must be a contiguous %s tensor" % (name, dtype))
    return C.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _prec(precision):
    if precision in ("fp64", FP64):
        return FP64
    if precision in ("fp32c", "fp32", FP32C):
        return FP32C
    raise ValueError("precision must be 'fp64' or 'fp32c'")


def cube_dtype(precision):
    torch = _torch()
    return torch.complex128 if _prec(precision) == FP64 else torch.complex64


def _vis_args(uu, vv, ww, vis_real, vis_img, weight):
    torch = _torch()
    return (_ptr(uu, torch.float64, "uu"), _ptr(vv, torch.float64, "vv"),
            _ptr(ww, torch.float64, "ww"), _ptr(vis_real, torch.float32, "vis_real"),
            _ptr(vis_img, torch.float32, "vis_img"), _ptr(weight, torch.float32, "weight"))


def _check_cube(cube, precision, nw, ny, nx):
    torch = _torch()
    want = cube_dtype(precision)
    if cube.dtype != want or cube.numel() < nw * ny * nx:
        raise TypeError("cube must be a %s tensor of at least %d cells" % (want, nw * ny * nx))
    return _ptr(cube, want, "cube")


def grid(uu, vv, ww, vis_real, vis_img, weight, num_w_planes, w_support, grid_size_x, grid_size_y,
         cube=None, precision="fp64", accumulate=False, dx=None, dw=None, freq_per_chan=1,
         polarizations=1):
    """Grid on the device; returns the cube [nw, ny, nx] (complex128 or complex64)."""
    torch = _torch()
    p = _prec(precision)
    n = uu.numel()
    if cube is None:
        cube = torch.empty((num_w_planes, grid_size_y, grid_size_x), dtype=cube_dtype(p),
                           device=uu.device)
        accumulate = False
    cp = _check_cube(cube, p, num_w_planes, grid_size_y, grid_size_x)
    check("lofar_grid_dev", lib().lofar_grid_dev(
        p, int(accumulate), num_w_planes, n, freq_per_chan, polarizations,
        *_vis_args(uu, vv, ww, vis_real, vis_img, weight),
        dx if dx is not None else 1.0 / grid_size_x, dw if dw is not None else 1.0 / num_w_planes,
        w_support, grid_size_x, grid_size_y, cp, _stream()))
    return cube


def fft(cube, sign=+1, scale=None, out=None, precision=None):
    """Per-plane 2D transform of a [nw, ny, nx] cube; sign=+1 is FFTW_BACKWARD, default scale 1/(nx*ny)."""
    torch = _torch()
    nw, ny, nx = cube.shape
    p = FP64 if cube.dtype == torch.complex128 else FP32C
    if precision is not None and _prec(precision) != p:
        raise TypeError("cube dtype does not match precision")
    if out is None:
        out = torch.empty_like(cube)
    check("lofar_fft_dev", lib().lofar_fft_dev(
        p, sign, _ptr(cube, cube.dtype, "cube"), _ptr(out, cube.dtype, "out"), nx, ny, nw,
        (1.0 / (nx * ny)) if scale is None else scale, _stream()))
    return out


def phase(cube, resolution, wmin, wmax, image_real=None, image_imag=None, accumulate=False,
          xaxistot=None, yaxistot=None):
    """phase_correction() on a device cube; returns (image_real, image_imag) float64 [ny, nx]."""
    torch = _torch()
    nw, ny, nx = cube.shape
    p = FP64 if cube.dtype == torch.complex128 else FP32C
    if image_real is None:
        image_real = torch.empty((ny, nx), dtype=torch.float64, device=cube.device)
        image_imag = torch.empty((ny, nx), dtype=torch.float64, device=cube.device)
        accumulate = False
    check("lofar_phase_dev", lib().lofar_phase_dev(
        p, _ptr(cube, cube.dtype, "cube"), _ptr(image_real, torch.float64, "image_real"),
        _ptr(image_imag, torch.float64, "image_imag"), nx, ny, nw, xaxistot or nx, yaxistot or ny,
        resolution, wmin, wmax, int(accumulate), _stream()))
    return image_real, image_imag


def image(uu, vv, ww, vis_real, vis_img, weight, num_w_planes, w_support, grid_size_x,
          grid_size_y, resolution, wmin, wmax, precision="fp64", cube=None, image_real=None,
          image_imag=None, accumulate=False, freq_per_chan=1, polarizations=1,
          static_workspace=False):
    """Whole hot path on the device: grid -> per-plane FFT -> phase correction + plane sum.
    ``static_workspace=True`` calls lofar_image_dev_static: the same pipeline with the
    library's workspace in its static per-device array instead of hipMalloc."""
    torch = _torch()
    p = _prec(precision)
    dev = uu.device
    if cube is None:
        cube = torch.empty((num_w_planes, grid_size_y, grid_size_x), dtype=cube_dtype(p), device=dev)
    cp = _check_cube(cube, p, num_w_planes, grid_size_y, grid_size_x)
    if image_real is None:
        image_real = torch.empty((grid_size_y, grid_size_x), dtype=torch.float64, device=dev)
        image_imag = torch.empty((grid_size_y, grid_size_x), dtype=torch.float64, device=dev)
        accumulate = False
    name = "lofar_image_dev_static" if static_workspace else "lofar_image_dev"
    check(name, getattr(lib(), name)(
        p, num_w_planes, uu.numel(), freq_per_chan, polarizations,
        *_vis_args(uu, vv, ww, vis_real, vis_img, weight),
        1.0 / grid_size_x, 1.0 / num_w_planes, w_support, grid_size_x, grid_size_y,
        resolution, wmin, wmax, cp, _ptr(image_real, torch.float64, "image_real"),
        _ptr(image_imag, torch.float64, "image_imag"), int(accumulate), _stream()))
    return image_real, image_imag


def phase_adjoint(image_real, image_imag, num_w_planes, resolution, wmin, wmax, precision="fp64",
                  cube=None):
    torch = _torch()
    p = _prec(precision)
    ny, nx = image_real.shape
    if cube is None:
        cube = torch.empty((num_w_planes, ny, nx), dtype=cube_dtype(p), device=image_real.device)
    cp = _check_cube(cube, p, num_w_planes, ny, nx)
    check("lofar_phase_adjoint_dev", lib().lofar_phase_adjoint_dev(
        p, _ptr(image_real, torch.float64, "image_real"),
        _ptr(image_imag, torch.float64, "image_imag"), cp, nx, ny, num_w_planes, nx, ny,
        resolution, wmin, wmax, _stream()))
    return cube


def _model_out(n, p, device, model_real, model_imag):
    torch = _torch()
    dt = torch.float64 if p == FP64 else torch.float32
    if model_real is None:
        model_real = torch.empty(n, dtype=dt, device=device)
        model_imag = torch.empty(n, dtype=dt, device=device)
    return model_real, model_imag, _ptr(model_real, dt, "model_real"), _ptr(model_imag, dt, "model_imag")


def degrid_cube(cube, uu, vv, ww, weight, w_support, model_real=None, model_imag=None,
                freq_per_chan=1, polarizations=1):
    """Gaussian interpolation of a device cube at every sample (adjoint of grid)."""
    torch = _torch()
    nw, ny, nx = cube.shape
    p = FP64 if cube.dtype == torch.complex128 else FP32C
    n = uu.numel()
    mr, mi, pr, pi = _model_out(n * freq_per_chan * polarizations, p, uu.device, model_real,
                                model_imag)
    check("lofar_degrid_cube_dev", lib().lofar_degrid_cube_dev(
        p, nw, n, freq_per_chan, polarizations, _ptr(uu, torch.float64, "uu"),
        _ptr(vv, torch.float64, "vv"), _ptr(ww, torch.float64, "ww"),
        _ptr(weight, torch.float32, "weight"), 1.0 / nx, 1.0 / nw, w_support, nx, ny,
        _ptr(cube, cube.dtype, "cube"), pr, pi, _stream()))
    return mr, mi


def degrid(image_real, image_imag, uu, vv, ww, weight, num_w_planes, w_support, resolution, wmin,
           wmax, precision="fp64", cube=None, model_real=None, model_imag=None):
    """Exact adjoint of ``image``: phase adjoint + forward FFT/(NxNy) + interpolation."""
    torch = _torch()
    p = _prec(precision)
    ny, nx = image_real.shape
    if cube is None:
        cube = torch.empty((num_w_planes, ny, nx), dtype=cube_dtype(p), device=image_real.device)
    cp = _check_cube(cube, p, num_w_planes, ny, nx)
    n = uu.numel()
    mr, mi, pr, pi = _model_out(n, p, uu.device, model_real, model_imag)
    check("lofar_degrid_dev", lib().lofar_degrid_dev(
        p, num_w_planes, n, 1, 1, _ptr(uu, torch.float64, "uu"), _ptr(vv, torch.float64, "vv"),
        _ptr(ww, torch.float64, "ww"), _ptr(weight, torch.float32, "weight"), 1.0 / nx,
        1.0 / num_w_planes, w_support, nx, ny, resolution, wmin, wmax,
        _ptr(image_real, torch.float64, "image_real"), _ptr(image_imag, torch.float64, "image_imag"),
        cp, pr, pi, _stream()))
    return mr, mi


class Plan:
    """One sort of fixed samples (lofar_plan_create) for repeated ``image`` and
    ``degrid`` calls with new visibilities or a new image: the same results as the
    unplanned calls, without re-sorting or host synchronisation.  The plan owns
    device memory (released by ``close`` or when the object is collected); the
    degridding weights are fixed at creation, the imaging weights come with each call."""

    def __init__(self, uu, vv, ww, weight, num_w_planes, w_support, grid_size_x, grid_size_y,
                 precision="fp64", freq_per_chan=1, polarizations=1):
        torch = _torch()
        self.precision = _prec(precision)
        self.nw, self.support = num_w_planes, w_support
        self.nx, self.ny = grid_size_x, grid_size_y
        self.n, self.F, self.P = uu.numel(), freq_per_chan, polarizations
        self.device = uu.device
        h = C.c_void_p()
        check("lofar_plan_create", lib().lofar_plan_create(
            self.precision, num_w_planes, self.n, freq_per_chan, polarizations,
            _ptr(uu, torch.float64, "uu"), _ptr(vv, torch.float64, "vv"),
            _ptr(ww, torch.float64, "ww"), _ptr(weight, torch.float32, "weight"),
            1.0 / grid_size_x, 1.0 / num_w_planes, w_support, grid_size_x, grid_size_y,
            _stream(), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lofar_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def image(self, vis_real, vis_img, weight, resolution, wmin, wmax, cube=None,
              image_real=None, image_imag=None, accumulate=False):
        """``image`` on the planned samples; returns (image_real, image_imag)."""
        torch = _torch()
        if cube is None:
            cube = torch.empty((self.nw, self.ny, self.nx), dtype=cube_dtype(self.precision),
                               device=self.device)
        cp = _check_cube(cube, self.precision, self.nw, self.ny, self.nx)
        if image_real is None:
            image_real = torch.empty((self.ny, self.nx), dtype=torch.float64, device=self.device)
            image_imag = torch.empty((self.ny, self.nx), dtype=torch.float64, device=self.device)
            accumulate = False
        check("lofar_plan_image", lib().lofar_plan_image(
            self._h, _ptr(vis_real, torch.float32, "vis_real"),
            _ptr(vis_img, torch.float32, "vis_img"), _ptr(weight, torch.float32, "weight"),
            resolution, wmin, wmax, cp, _ptr(image_real, torch.float64, "image_real"),
            _ptr(image_imag, torch.float64, "image_imag"), int(accumulate), _stream()))
        return image_real, image_imag

    def degrid(self, image_real, image_imag, resolution, wmin, wmax, cube=None, model_real=None,
               model_imag=None):
        """``degrid`` on the planned samples; returns (model_real, model_imag)."""
        torch = _torch()
        if cube is None:
            cube = torch.empty((self.nw, self.ny, self.nx), dtype=cube_dtype(self.precision),
                               device=self.device)
        cp = _check_cube(cube, self.precision, self.nw, self.ny, self.nx)
        mr, mi, pr, pi = _model_out(self.n * self.F * self.P, self.precision, self.device,
                                    model_real, model_imag)
        check("lofar_plan_degrid", lib().lofar_plan_degrid(
            self._h, resolution, wmin, wmax, _ptr(image_real, torch.float64, "image_real"),
            _ptr(image_imag, torch.float64, "image_imag"), cp, pr, pi, _stream()))
        return mr, mi


# ---------------------------------------------------------------------------
# per-stage kernel timing (HIP events on the stream the kernels run on)

def profile_enable(on=True):
    lib().lofar_profile_enable(int(on))


def profile_read():
    """Returns {stage_name: (total_ms, launches)} and clears the record."""
    tot = (C.c_double * 10)()
    cnt = (C.c_int * 10)()
    check("lofar_profile_read", lib().lofar_profile_read(tot, cnt))
    return {lib().lofar_stage_name(i).decode(): (tot[i], cnt[i]) for i in range(10) if cnt[i]}
