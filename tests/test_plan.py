"""Plans (lofar_plan_*): one sort of fixed samples reused by imaging and degridding.

A plan's calls must give the unplanned entry points' results bit for bit: the tiles sum
exact fixed-point integers (independent of the entries' order) and the degridder computes
every sample on its own, so the sort's order cannot show.  The unplanned paths are pinned
against the oracle in test_gpu_parity.py; here the oracle checks one planned image and one
planned degridding as well.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import H, T, dev_inputs, random_case, rel, rel_nan, synth

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
@pytest.mark.parametrize("nx,ny,support", [(256, 256, 7), (300, 300, 7), (192, 160, 9),
                                           (256, 256, 21)])
def test_plan_image_and_degrid_match_unplanned_bitwise(cuda, precision, nx, ny, support):
    """Imaging two visibility sets and degridding two images through one plan gives the
    bits of lofar_image_dev and lofar_degrid_dev on the same inputs: tile sorts at powers
    of two and 300x300 (mixed radix); the unsorted path of a 21-cell window (its images to
    rounding: global atomic adds)."""
    from lofar_amd import imager
    nw = 3
    d = random_case(71, 40000, nx, ny, nw, edge=True)
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    res, wmin, wmax = 1.0 / nx, -40.0, 40.0
    plan = imager.Plan(uu, vv, ww, wt, nw, support, nx, ny, precision=precision)
    rng = np.random.default_rng(72)
    for k in range(2):
        vr2 = T(rng.standard_normal(40000).astype(np.float32), cuda)
        vi2 = T(rng.standard_normal(40000).astype(np.float32), cuda)
        want = imager.image(uu, vv, ww, vr2, vi2, wt, nw, support, nx, ny, res, wmin, wmax,
                            precision=precision)
        got = plan.image(vr2, vi2, wt, res, wmin, wmax)
        if support <= 17:
            assert _bits_equal(H(got[0]), H(want[0])) and _bits_equal(H(got[1]), H(want[1])), k
        else:  # the unsorted gridder's global atomic adds sum in a varying order
            # (pixels with l^2 + m^2 > 1 are NaN in both, as in the reference)
            e = max(rel_nan(H(got[0]), H(want[0])), rel_nan(H(got[1]), H(want[1])))
            assert e < (1e-12 if precision == "fp64" else 1e-6), (k, e)
        img_re = T(rng.standard_normal((ny, nx)), cuda)
        img_im = T(rng.standard_normal((ny, nx)), cuda)
        mw = imager.degrid(img_re, img_im, uu, vv, ww, wt, nw, support, res, wmin, wmax,
                           precision=precision)
        mg = plan.degrid(img_re, img_im, res, wmin, wmax)
        assert _bits_equal(H(mg[0]), H(mw[0])) and _bits_equal(H(mg[1]), H(mw[1])), k
    plan.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32c", 1e-6)])
def test_plan_matches_oracle_and_outlives_the_callers_arrays(cuda, precision, tol):
    """The plan keeps what it needs: after the caller overwrites uu, vv, ww and the
    weights, the planned image and degridding still match the oracle on the original
    samples; accumulate mode adds to the image."""
    from lofar_amd import imager
    n, nx, nw = 60000, 256, 4
    d = synth(n, nx, nw, seed=73)
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    plan = imager.Plan(uu, vv, ww, wt, nw, 7, nx, nx, precision=precision)
    for t in (uu, vv, ww):
        t.fill_(0.5)
    wt.fill_(3.0)
    wt_img = T(d["weight"], cuda)
    want = O.image(nw, d["uu"], d["vv"], d["ww"], d["vis_real"], d["vis_img"], d["weight"], 7,
                   nx, nx, d["resolution"], d["wmin"], d["wmax"])
    r, i = plan.image(vr, vi, wt_img, d["resolution"], d["wmin"], d["wmax"])
    assert rel(H(r) + 1j * H(i), want) < tol
    plan.image(vr, vi, wt_img, d["resolution"], d["wmin"], d["wmax"], image_real=r,
               image_imag=i, accumulate=True)
    assert rel(H(r) + 1j * H(i), 2 * want) < tol
    rng = np.random.default_rng(74)
    y = rng.standard_normal((nx, nx)) + 1j * rng.standard_normal((nx, nx))
    want_m = O.degrid_image(y, nw, d["uu"], d["vv"], d["ww"], d["weight"], 7, d["resolution"],
                            d["wmin"], d["wmax"])
    mr, mi = plan.degrid(T(y.real.copy(), cuda), T(y.imag.copy(), cuda), d["resolution"],
                         d["wmin"], d["wmax"])
    assert rel(H(mr) + 1j * H(mi), want_m) < (1e-10 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_plan_channels_and_polarisations(cuda, precision):
    """Two channels and two polarisations: the planned image sums them with the call's
    weights as lofar_image_dev does (bitwise; in FP32C through the per-sample sums the packed
    gridder gathers), and the planned degridding writes every (channel, polarisation) with
    the plan's weights as lofar_degrid_dev does."""
    import ctypes as C
    import torch
    from lofar_amd import imager
    from lofar_amd._lib import FP32C, FP64, check, lib
    n, nx, nw, F, P = 20000, 128, 2, 2, 2
    d = random_case(75, n, nx, nx, nw, F=F, P=P, edge=True)
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    plan = imager.Plan(uu, vv, ww, wt, nw, 7, nx, nx, precision=precision, freq_per_chan=F,
                       polarizations=P)
    args = (nw, 7, nx, nx, 1.0 / nx, -10.0, 10.0)
    want = imager.image(uu, vv, ww, vr, vi, wt, *args, precision=precision, freq_per_chan=F,
                        polarizations=P)
    got = plan.image(vr, vi, wt, 1.0 / nx, -10.0, 10.0)
    assert _bits_equal(H(got[0]), H(want[0])) and _bits_equal(H(got[1]), H(want[1]))
    fp64 = precision == "fp64"
    img = T(np.random.default_rng(76).standard_normal((nx, nx)), cuda)
    zero = torch.zeros_like(img)
    cube = torch.empty((nw, nx, nx), dtype=torch.complex128 if fp64 else torch.complex64,
                       device=cuda)
    mr = torch.empty(n * F * P, dtype=torch.float64 if fp64 else torch.float32, device=cuda)
    mi = torch.empty_like(mr)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr())
    check("lofar_degrid_dev", lib().lofar_degrid_dev(
        FP64 if fp64 else FP32C, nw, n, F, P, p(uu), p(vv), p(ww), p(wt), 1.0 / nx, 1.0 / nw, 7,
        nx, nx, 1.0 / nx, -10.0, 10.0, p(img), p(zero), p(cube), p(mr), p(mi), s))
    gr, gi = plan.degrid(img, zero, 1.0 / nx, -10.0, 10.0)
    assert gr.dtype == mr.dtype
    assert _bits_equal(H(gr), H(mr)) and _bits_equal(H(gi), H(mi))


def test_plan_rejects_bad_samples_and_null_plans(cuda):
    """A sample outside the w-planes fails lofar_plan_create with LOFAR_ECOORD (no plan);
    a null plan is LOFAR_EARG."""
    import ctypes as C
    from lofar_amd import imager
    from lofar_amd._lib import EARG, ECOORD, LofarError, lib
    d = random_case(77, 1000, 64, 64, 2, edge=False)
    d["ww"] = d["ww"].copy()
    d["ww"][10] = 2.5
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    with pytest.raises(LofarError) as e:
        imager.Plan(uu, vv, ww, wt, 2, 7, 64, 64, precision="fp32c")
    assert e.value.code == ECOORD
    assert lib().lofar_plan_image(None, None, None, None, 1.0, 0.0, 1.0, None, None, None, 0,
                                  None) == EARG
    assert lib().lofar_plan_destroy(None) == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_operator_with_and_without_plan_agree_bitwise(cuda, precision):
    """ImagingOperator's forward and adjoint (the chi^2 fit's and major cycles' calls) give
    the same bits with the plan (default) and without it."""
    import torch
    from lofar_amd.autograd import ImagingOperator
    n, nx, nw = 50000, 256, 4
    d = synth(n, nx, nw, seed=78)
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    args = (uu, vv, ww, wt, nw, 7, nx, nx, d["resolution"], d["wmin"], d["wmax"])
    a = ImagingOperator(*args, precision=precision)
    b = ImagingOperator(*args, precision=precision, plan=False)
    g = torch.Generator(device=cuda)
    g.manual_seed(79)
    x_re = torch.randn((nx, nx), generator=g, dtype=torch.float64, device=cuda)
    x_im = torch.randn((nx, nx), generator=g, dtype=torch.float64, device=cuda)
    fa, fb = a.forward(x_re, x_im), b.forward(x_re, x_im)
    assert _bits_equal(H(fa[0]), H(fb[0])) and _bits_equal(H(fa[1]), H(fb[1]))
    ya, yb = a.adjoint(vr.double(), vi.double()), b.adjoint(vr.double(), vi.double())
    assert _bits_equal(H(ya[0]), H(yb[0])) and _bits_equal(H(ya[1]), H(yb[1]))
    assert a._plan is not None and b._plan is None


@pytest.mark.parametrize("case", ["bright_outlier", "dense_block"])
def test_plan_fp32c_fallback_tiles_match_unplanned(cuda, case):
    """The packed FP32C gridder's fallbacks gather the plan's per-sample sums too: a tile with
    one 4x4-cell block under 600 windows (two-word fixed-point halves: exact sums) gives the
    unplanned image's bits; a tile whose brightest sample is 1e6 times the rest sums in double
    with LDS atomics, whose order is not fixed, so it agrees to FP32C's tolerance."""
    from lofar_amd import imager
    n, nx, nw = 20000, 256, 2
    d = random_case(81, n, nx, nx, nw, edge=False, nan=0.0)
    rng = np.random.default_rng(82)
    if case == "bright_outlier":
        d["vis_real"] = d["vis_real"].copy()
        d["vis_real"][123] = 1e6
    else:
        # 600 samples inside one 4x4-cell block of one plane
        k = np.arange(600)
        d["uu"] = d["uu"].copy()
        d["vv"] = d["vv"].copy()
        d["ww"] = d["ww"].copy()
        d["uu"][k] = (100.0 + rng.uniform(0.0, 4.0, k.size)) / nx
        d["vv"][k] = (60.0 + rng.uniform(0.0, 4.0, k.size)) / nx
        d["ww"][k] = 0.25
    uu, vv, ww, vr, vi, wt = dev_inputs(d, cuda)
    plan = imager.Plan(uu, vv, ww, wt, nw, 7, nx, nx, precision="fp32c")
    args = (nw, 7, nx, nx, 1.0 / nx, -10.0, 10.0)
    want = imager.image(uu, vv, ww, vr, vi, wt, *args, precision="fp32c")
    got = plan.image(vr, vi, wt, 1.0 / nx, -10.0, 10.0)
    if case == "dense_block":
        assert _bits_equal(H(got[0]), H(want[0])) and _bits_equal(H(got[1]), H(want[1]))
    else:
        e = max(rel_nan(H(got[0]), H(want[0])), rel_nan(H(got[1]), H(want[1])))
        assert e < 1e-6, e
    plan.close()
