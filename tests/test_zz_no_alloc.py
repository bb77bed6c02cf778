"""lofar_image_dev_static, the allocation-free variant of lofar_image_dev: through the C
ABI it gives lofar_image_dev's image (FP64 within 1e-12, FP32C within 1e-6, relative L2)
on two calls in a row, the second with fewer samples on a smaller grid (the static
workspace keeps the first call's contents), and a call needing more workspace than the
static array holds fails with LOFAR_ESCRATCH and leaves the image untouched."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(name, d, nw, support, nx, precision, cube, re, im):
    import torch
    from lofar_amd._lib import lib
    p = 0 if precision == "fp64" else 1
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    return getattr(lib(), name)(
        p, nw, d["uu"].numel(), 1, 1, ptr(d["uu"]), ptr(d["vv"]), ptr(d["ww"]),
        ptr(d["vis_real"]), ptr(d["vis_img"]), ptr(d["weight"]), 1.0 / nx, 1.0 / nw, support,
        nx, nx, d["resolution"], d["wmin"], d["wmax"], ptr(cube), ptr(re), ptr(im), 0,
        C.c_void_p(torch.cuda.current_stream().cuda_stream))


def _image(name, d, nw, support, nx, precision, dev):
    import torch
    from lofar_amd import imager
    cube = torch.full((nw, nx, nx), float("nan"), dtype=imager.cube_dtype(precision), device=dev)
    re = torch.full((nx, nx), float("nan"), dtype=torch.float64, device=dev)
    im = torch.full_like(re, float("nan"))
    st = _call(name, d, nw, support, nx, precision, cube, re, im)
    torch.cuda.synchronize()
    return st, re.cpu().numpy() + 1j * im.cpu().numpy()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-6)])
def test_static_workspace_matches_lofar_image_dev(cuda, precision, tol):
    from lofar_amd.synth import make_obs
    for n, nx, nw, support in ((400000, 1024, 8, 7), (30000, 256, 4, 9)):
        d = make_obs(n, nx, nw, seed=n + nx, backend="torch", device=cuda)
        st0, want = _image("lofar_image_dev", d, nw, support, nx, precision, cuda)
        st1, got = _image("lofar_image_dev_static", d, nw, support, nx, precision, cuda)
        assert st0 == 0 and st1 == 0, (st0, st1)
        err = float(np.linalg.norm(got - want) / np.linalg.norm(want))
        print("static workspace %s n=%d nx=%d: rel %.3e" % (precision, n, nx, err))
        assert np.isfinite(got).all() and err <= tol


def test_static_workspace_rejects_oversize_call(cuda):
    import torch
    from lofar_amd._lib import ESCRATCH
    from lofar_amd.synth import make_obs
    # 8M FP64 samples need ~0.5 GB of bucket records and entries: beyond the 256 MiB array
    n, nx, nw = 8 * 1024 * 1024, 1024, 8
    d = make_obs(n, nx, nw, seed=3, backend="torch", device=cuda)
    cube = torch.empty((nw, nx, nx), dtype=torch.complex128, device=cuda)
    re = torch.full((nx, nx), 7.0, dtype=torch.float64, device=cuda)
    im = torch.full_like(re, -7.0)
    st = _call("lofar_image_dev_static", d, nw, 7, nx, "fp64", cube, re, im)
    torch.cuda.synchronize()
    assert st == ESCRATCH
    assert bool((re == 7.0).all()) and bool((im == -7.0).all())
