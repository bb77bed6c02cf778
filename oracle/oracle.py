"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  It loads

* ``liboracle.so``: the clean-room C restatement in ``oracle.c`` of wstack()
  (w-stacking_omp.c:25-145), the FFTW_BACKWARD/(Nx*Ny) plane transform
  (fourier_transform.c:7-74) and phase_correction() (phase_correction.cu:89-210);
* ``_ref/*.so`` when present: the reference's own w-stacking_omp.c,
  w-stacking.cu and phase_correction.cu compiled unchanged as C by
  ``oracle/Makefile``; these pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)

OK, EARG, ECOORD = 0, -1, -2


def _load(path):
    return C.CDLL(path) if os.path.exists(path) else None


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.oracle_wstack.restype = C.c_int
        L.oracle_wstack.argtypes = [C.c_int, C.c_long, C.c_long, C.c_long, _dp, _dp, _dp, _fp, _fp,
                                    _fp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp]
        L.oracle_fft_planes.restype = C.c_int
        L.oracle_fft_planes.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.oracle_phase_correction.restype = None
        L.oracle_phase_correction.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.oracle_phase_adjoint.restype = None
        L.oracle_phase_adjoint.argtypes = L.oracle_phase_correction.argtypes
        L.oracle_degrid.restype = C.c_int
        L.oracle_degrid.argtypes = [C.c_int, C.c_long, C.c_long, C.c_long, _dp, _dp, _dp, _fp,
                                    C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.oracle_resolution.restype = C.c_double
        L.oracle_resolution.argtypes = [C.c_double, C.c_double]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def _c64(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def wstack(nw, uu, vv, ww, vis_real, vis_img, weight, dx, dw, support, nx, ny, grid=None,
           freq_per_chan=1, polarizations=1):
    """Oracle gridding; returns (status, grid[nw, ny, nx] complex128), accumulating into grid."""
    uu, vv, ww = (_c64(a, np.float64) for a in (uu, vv, ww))
    vr, vi, wt = (_c64(a, np.float32) for a in (vis_real, vis_img, weight))
    if grid is None:
        grid = np.zeros((nw, ny, nx), np.complex128)
    g = grid.view(np.float64)
    st = lib().oracle_wstack(nw, len(uu), freq_per_chan, polarizations, _d(uu), _d(vv), _d(ww),
                             _f(vr), _f(vi), _f(wt), dx, dw, support, nx, ny, _d(g))
    return st, grid


def fft_planes(cube, sign=+1, scale=None):
    """Oracle per-plane 2D transform; sign=+1 is FFTW_BACKWARD.  Default scale 1/(nx*ny)."""
    cube = np.ascontiguousarray(cube, np.complex128)
    nw, ny, nx = cube.shape
    if scale is None:
        scale = 1.0 / (nx * ny)
    out = np.empty_like(cube)
    st = lib().oracle_fft_planes(_d(cube.view(np.float64)), _d(out.view(np.float64)), nx, ny, nw,
                                 sign, scale)
    if st != OK:
        raise ValueError("oracle_fft_planes failed: %d" % st)
    return out


def phase_correction(cube, resolution, wmin, wmax, xtot=None, ytot=None, phase_on=True, image=None):
    """Oracle phase correction + plane sum; returns the complex image (accumulated)."""
    cube = np.ascontiguousarray(cube, np.complex128)
    nw, ny, nx = cube.shape
    if image is None:
        image = np.zeros((ny, nx), np.complex128)
    re = np.ascontiguousarray(image.real)
    im = np.ascontiguousarray(image.imag)
    lib().oracle_phase_correction(_d(cube.view(np.float64)), _d(re), _d(im), nx, ny, nw,
                                  xtot or nx, ytot or ny, resolution, wmin, wmax, int(phase_on))
    return re + 1j * im


def phase_adjoint(image, nw, resolution, wmin, wmax, phase_on=True):
    image = np.asarray(image, np.complex128)
    ny, nx = image.shape
    re = np.ascontiguousarray(image.real)
    im = np.ascontiguousarray(image.imag)
    out = np.empty((nw, ny, nx), np.complex128)
    lib().oracle_phase_adjoint(_d(re), _d(im), _d(out.view(np.float64)), nx, ny, nw, nx, ny,
                               resolution, wmin, wmax, int(phase_on))
    return out


def degrid(cube, uu, vv, ww, weight, dx, dw, support, freq_per_chan=1, polarizations=1):
    cube = np.ascontiguousarray(cube, np.complex128)
    nw, ny, nx = cube.shape
    uu, vv, ww = (_c64(a, np.float64) for a in (uu, vv, ww))
    wt = _c64(weight, np.float32)
    n = len(uu)
    mr = np.zeros(n * freq_per_chan * polarizations)
    mi = np.zeros_like(mr)
    st = lib().oracle_degrid(nw, n, freq_per_chan, polarizations, _d(uu), _d(vv), _d(ww), _f(wt),
                             dx, dw, support, nx, ny, _d(cube.view(np.float64)), _d(mr), _d(mi))
    return st, mr + 1j * mi


def image(nw, uu, vv, ww, vis_real, vis_img, weight, support, nx, ny, resolution, wmin, wmax):
    """Whole hot path on the CPU: wstack -> per-plane FFT -> phase correction."""
    st, cube = wstack(nw, uu, vv, ww, vis_real, vis_img, weight, 1.0 / nx, 1.0 / nw, support, nx, ny)
    if st != OK:
        raise ValueError("oracle wstack failed: %d" % st)
    return phase_correction(fft_planes(cube), resolution, wmin, wmax)


def degrid_image(img, nw, uu, vv, ww, weight, support, resolution, wmin, wmax):
    """Exact adjoint of image(): phase adjoint -> forward FFT/(Nx*Ny) -> interpolation."""
    ny, nx = img.shape
    planes = phase_adjoint(img, nw, resolution, wmin, wmax)
    planes = fft_planes(planes, sign=-1, scale=1.0 / (nx * ny))
    st, model = degrid(planes, uu, vv, ww, weight, 1.0 / nx, 1.0 / nw, support)
    if st != OK:
        raise ValueError("oracle degrid failed: %d" % st)
    return model


def resolution(uvmin, uvmax):
    return lib().oracle_resolution(uvmin, uvmax)


# --------------------------------------------------------------------------
# the reference's own C, compiled in place (oracle/_ref); None when absent

class RefLib:
    """The reference's wstack()/phase_correction() built from its sources."""

    def __init__(self, fast=False):
        d = os.path.join(_HERE, "_ref", "fast" if fast else "")
        self.serial = _load(os.path.join(d, "libref_wstack_serial.so"))
        self.omp = _load(os.path.join(d, "libref_wstack_omp.so"))
        self.phase = _load(os.path.join(d, "libref_phase.so"))
        for L in (self.serial, self.omp):
            if L is not None:
                L.wstack.restype = None
                L.wstack.argtypes = [C.c_int, C.c_long, C.c_long, C.c_long, _dp, _dp, _dp, _fp,
                                     _fp, _fp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                     _dp, C.c_int]
                L.test.restype = C.c_int
                L.test.argtypes = [C.c_int]
        if self.phase is not None:
            self.phase.phase_correction.restype = None
            self.phase.phase_correction.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int,
                                                    C.c_int, C.c_int, C.c_double, C.c_double,
                                                    C.c_double, C.c_int]

    @property
    def available(self):
        return self.serial is not None and self.omp is not None and self.phase is not None

    def wstack(self, nw, uu, vv, ww, vis_real, vis_img, weight, dx, dw, support, nx, ny,
               grid=None, freq_per_chan=1, polarizations=1, variant="serial", num_threads=1):
        L = self.serial if variant == "serial" else self.omp
        uu, vv, ww = (_c64(a, np.float64) for a in (uu, vv, ww))
        vr, vi, wt = (_c64(a, np.float32) for a in (vis_real, vis_img, weight))
        if grid is None:
            grid = np.zeros((nw, ny, nx), np.complex128)
        L.wstack(nw, len(uu), freq_per_chan, polarizations, _d(uu), _d(vv), _d(ww), _f(vr),
                 _f(vi), _f(wt), dx, dw, support, nx, ny, _d(grid.view(np.float64)), num_threads)
        return grid

    def phase_correction(self, cube, resolution, wmin, wmax, num_threads=1, image=None):
        cube = np.ascontiguousarray(cube, np.complex128)
        nw, ny, nx = cube.shape
        if image is None:
            image = np.zeros((ny, nx), np.complex128)
        re = np.ascontiguousarray(image.real)
        im = np.ascontiguousarray(image.imag)
        self.phase.phase_correction(_d(cube.view(np.float64)), _d(re), _d(im), nx, ny, nw, nx, ny,
                                    resolution, wmin, wmax, num_threads)
        return re + 1j * im
