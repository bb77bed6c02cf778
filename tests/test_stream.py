"""Multi-channel streaming (lofar_amd.stream): per-channel uvw scaling and normalisation,
chunked host->device streaming, imaging of the streamed samples against the oracle, and the
major-cycle loop."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_stream_expands_channels_exactly(cuda):
    import torch
    from lofar_amd.stream import C_LIGHT, ONE_MINUS, make_multichannel
    obs = make_multichannel(5003, 256, nchan=4, seed=51, device=cuda, gen_chunk=1000)
    out, secs, nbytes = obs.to_device(cuda, chunk_samples=4 * 777)
    base = obs.base_uvw.numpy()
    scale = (np.asarray(obs.freqs) / C_LIGHT).reshape(1, -1)
    want_u = ((base[:, 0:1] * scale) / obs.span + 0.5).reshape(-1)
    want_w = np.clip(((base[:, 2:3] * scale) - obs.wmin) * (1.0 / (obs.wmax - obs.wmin)), 0.0,
                     ONE_MINUS).reshape(-1)
    assert np.array_equal(out["uu"].cpu().numpy(), want_u)
    assert np.array_equal(out["ww"].cpu().numpy(), want_w)
    assert np.array_equal(out["vis_real"].cpu().numpy(), obs.vis_re.numpy().reshape(-1))
    assert out["uu"].min() > 0 and out["uu"].max() < 1 and nbytes == 5003 * (24 + 32)
    torch.cuda.synchronize()


def test_streamed_samples_image_matches_oracle(cuda):
    from lofar_amd import imager
    from lofar_amd.stream import make_multichannel
    obs = make_multichannel(8000, 256, nchan=8, seed=52, device=cuda)
    s, _, _ = obs.to_device(cuda, chunk_samples=10000)
    res = 1.0 / int(obs.span)
    h = {k: v.cpu().numpy() for k, v in s.items()}
    want = O.image(8, h["uu"], h["vv"], h["ww"], h["vis_real"], h["vis_img"], h["weight"], 9, 256,
                   256, res, obs.wmin, obs.wmax)
    for prec, tol in (("fp64", 1e-10), ("fp32c", 1e-6)):
        r, i = imager.image(s["uu"], s["vv"], s["ww"], s["vis_real"], s["vis_img"], s["weight"], 8,
                            9, 256, 256, res, obs.wmin, obs.wmax, precision=prec)
        got = r.cpu().numpy() + 1j * i.cpu().numpy()
        assert np.linalg.norm(got - want) < tol * np.linalg.norm(want)


def test_major_cycles_reduce_chi2(cuda):
    import torch
    from lofar_amd.autograd import ImagingOperator
    from lofar_amd.stream import major_cycles, make_multichannel
    obs = make_multichannel(20000, 256, nchan=8, seed=53, device=cuda)
    s, _, _ = obs.to_device(cuda)
    op = ImagingOperator(s["uu"], s["vv"], s["ww"], None, 8, 9, 256, 256, 1.0 / int(obs.span),
                         obs.wmin, obs.wmax, precision="fp32c")
    x0 = torch.zeros((256, 256), dtype=torch.float64, device=cuda)
    hist = []
    major_cycles(op, s["vis_real"], s["vis_img"], s["weight"], x0, x0.clone(), cycles=5,
                 history=hist)
    h = np.array([float(v) for v in hist])
    assert np.all(np.diff(h) <= 1e-6 * h[0]) and h[-1] < 0.95 * h[0]
