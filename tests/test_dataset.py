"""The reference's file formats (lofar_amd.dataset): paramfile, meta.txt, binMS arrays,
the one-rank sector selection, and the output dumps.  CPU tests write a dataset in the
reference's layout and read it back; the GPU test runs the whole driver on it and checks
the image file against the oracle pipeline on the same sector samples."""
import os

import numpy as np
import pytest

from lofar_amd import dataset as D

PARAMS = """ndatasets              2
Datapath1              {p1}
Datapath2              {p2}
num_threads            2
w_support              7
grid_size_x            256
grid_size_y            256
num_w_planes           4
ufile\t\t       ucoord.bin
vfile    \t       vcoord.bin
wfile                  wcoord.bin
weightsfile\t       weights.bin
visrealfile            visibilities_real.bin
visimgfile             visibilities_img.bin
metafile               meta.txt
outfile2\t       grid_real.bin
outfile3 \t       grid_img.bin
fftfile2\t       fft_real.bin
fftfile3\t       fft_img.bin
verbose_level\t       1
"""


def _make(tmp, name, n, F, P, seed, nx=256):
    from lofar_amd.synth import make_obs
    d = make_obs(n, nx, 4, seed=seed)
    rng = np.random.default_rng(seed)
    vr = np.repeat(d["vis_real"], F * P).astype(np.float32) * rng.uniform(0.5, 1.5, n * F * P).astype(np.float32)
    vi = np.repeat(d["vis_img"], F * P).astype(np.float32)
    data = dict(uu=d["uu"], vv=d["vv"].copy(), ww=d["ww"], weight=np.ones(n * P, np.float32),
                vis_real=vr, vis_img=vi)
    # a few samples outside sector 0 ((int)vv == 1) and one with a negative vv
    data["vv"][5] = 1.25
    data["vv"][17] = -0.25 / nx
    meta = D.Meta(n, n * F * P, F, P, 10, 1.5, 2.0, 7, d["uvmin"], d["uvmax"], d["wmin"], d["wmax"])
    path = os.path.join(str(tmp), name) + os.sep
    D.write_dataset(path, meta, data)
    return path, meta, data, d


def test_paramfile_parsing_and_quirks(tmp_path):
    pf = tmp_path / "paramfile.txt"
    pf.write_text(PARAMS.format(p1="/a/", p2="/b/"))
    p = D.read_paramfile(str(pf))
    assert (p.ndatasets, p.num_threads, p.w_support) == (2, 2, 7)
    assert (p.grid_size_x, p.grid_size_y, p.num_w_planes) == (256, 256, 4)
    assert p.datapath_multi[:2] == ["/a/", "/b/"]
    assert p.files["wfile"] == "wcoord.bin" and p.outparam["fftfile3"] == "fft_img.bin"
    assert p.verbose_level == 0  # atoi of the key, as init.c:240 does
    # DatapathN before ndatasets > 1 is not recognised (init.c:243-252)
    pf.write_text("Datapath1 /a/\nDatapath2 /b/\nndatasets 2\n")
    p = D.read_paramfile(str(pf))
    assert p.datapath_multi[:2] == ["/a/", ""]
    assert D.out_name(p, 1, "fftfile2", "/o") == os.path.join("/o", "1fftfile2")


def test_meta_and_binms_round_trip(tmp_path):
    path, meta, data, _ = _make(tmp_path, "ds0", 3000, 2, 2, 41)
    p = D.read_paramfile(_paramfile(tmp_path, path, path))
    m0 = D.read_meta(os.path.join(path, "meta.txt"))
    assert (m0.Nmeasures, m0.freq_per_chan, m0.polarisations, m0.baselines) == (3000, 2, 2, 7)
    assert (m0.uvmin, m0.uvmax, m0.wmin, m0.wmax) == (meta.uvmin, meta.uvmax, meta.wmin, meta.wmax)
    m = D.metadata_calculation(m0)
    assert (m.Nmeasures, m.Nvis, m.Nweights) == (2000, 8000, 4000)
    got = D.read_data(path, p, m)
    for k in ("uu", "vv", "ww"):
        assert got[k].dtype == np.float64 and np.array_equal(got[k], data[k][:2000])
    assert np.array_equal(got["weight"], data["weight"][:4000])
    assert np.array_equal(got["vis_real"], data["vis_real"][:8000])
    assert np.array_equal(got["vis_img"], data["vis_img"][:8000])


def test_short_file_reads_as_zeros(tmp_path, capsys):
    path, meta, data, _ = _make(tmp_path, "ds0", 1500, 1, 1, 42)
    np.asarray(data["uu"][:100]).tofile(os.path.join(path, "ucoord.bin"))
    p = D.read_paramfile(_paramfile(tmp_path, path, path))
    got = D.read_data(path, p, D.metadata_calculation(D.read_meta(os.path.join(path, "meta.txt"))))
    assert np.array_equal(got["uu"][:100], data["uu"][:100]) and not got["uu"][100:].any()
    assert "error occurred while reading" in capsys.readouterr().out


def test_sector_selection_one_rank():
    rng = np.random.default_rng(3)
    n, F, P = 50, 2, 3
    vv = rng.uniform(0, 1, n)
    vv[[3, 9]] = [1.0, 1.7]     # (int)vv == 1: sector 1, never gridded by one rank
    vv[11] = -0.4               # (int) truncates toward zero: sector 0
    data = dict(uu=rng.uniform(0, 1, n), vv=vv, ww=rng.uniform(0, 1, n),
                weight=np.arange(n * P, dtype=np.float32),
                vis_real=np.arange(n * F * P, dtype=np.float32),
                vis_img=-np.arange(n * F * P, dtype=np.float32))
    meta = D.Meta(n, n * F * P, F, P, 0, 0, 0, 0, 0, 1, 0, 1)
    s = D.sector_samples(data, meta)
    keep = [i for i in range(n) if i not in (3, 9)][::-1]
    assert np.array_equal(s["uu"], data["uu"][keep]) and np.array_equal(s["vv"], vv[keep])
    assert np.array_equal(s["weight"].reshape(-1, P)[:, 0], np.array(keep, np.float32) * P)
    assert np.array_equal(s["vis_real"].reshape(-1, F * P)[:, 1], np.array(keep) * F * P + 1.0)


def test_output_dump_layouts(tmp_path):
    cube = (np.arange(24).reshape(2, 3, 4) + 1j * -np.arange(24).reshape(2, 3, 4)).astype(np.complex128)
    D.write_cube_dump(cube, str(tmp_path / "r.bin"), str(tmp_path / "i.bin"))
    assert np.array_equal(np.fromfile(tmp_path / "r.bin"), np.arange(24.0))
    assert np.array_equal(np.fromfile(tmp_path / "i.bin"), -np.arange(24.0))
    D.write_image(np.ones((3, 4)), np.zeros((3, 4)), str(tmp_path / "a.bin"), str(tmp_path / "b.bin"))
    assert os.path.getsize(tmp_path / "a.bin") == 96


def _paramfile(tmp, p1, p2):
    pf = os.path.join(str(tmp), "paramfile.txt")
    with open(pf, "w") as f:
        f.write(PARAMS.format(p1=p1, p2=p2))
    return pf


@pytest.mark.gpu
def test_driver_runs_reference_layout_on_device(cuda, tmp_path):
    from oracle import oracle as O
    from lofar_amd.imager import resolution_from_meta
    p1, m1, d1, _ = _make(tmp_path, "ds0", 6000, 1, 1, 43)
    p2, m2, d2, _ = _make(tmp_path, "ds1", 4000, 2, 2, 44)
    pf = _paramfile(tmp_path, p1, p2)
    out = tmp_path / "out"
    out.mkdir()
    imgs = D.run(pf, "fp64", str(out), write_data=True)
    p = D.read_paramfile(pf)
    for idx, (path, img) in enumerate(zip((p1, p2), imgs)):
        meta = D.metadata_calculation(D.read_meta(os.path.join(path, "meta.txt")))
        s = D.sector_samples(D.read_data(path, p, meta), meta)
        st, cube = O.wstack(4, s["uu"], s["vv"], s["ww"], s["vis_real"], s["vis_img"], s["weight"],
                            1 / 256, 1 / 4, 7, 256, 256, freq_per_chan=meta.freq_per_chan,
                            polarizations=meta.polarisations)
        assert st == O.OK
        got = np.fromfile(out / ("%dgrid_real.bin" % idx)) + 1j * np.fromfile(out / ("%dgrid_img.bin" % idx))
        assert np.linalg.norm(got - cube.reshape(-1)) < 1e-12 * np.linalg.norm(cube)
        want = O.phase_correction(O.fft_planes(cube), resolution_from_meta(meta.uvmin, meta.uvmax),
                                  meta.wmin, meta.wmax)
        re = np.fromfile(out / ("%dfft_real.bin" % idx)).reshape(256, 256)
        im = np.fromfile(out / ("%dfft_img.bin" % idx)).reshape(256, 256)
        assert np.linalg.norm(re + 1j * im - want) < 1e-10 * np.linalg.norm(want)
        assert np.array_equal(re, img[0])
    # the fused path (no cube dumps) gives the same image
    imgs2 = D.run(pf, "fp64", str(out), write_data=False)
    for a, b in zip(imgs, imgs2):
        assert np.linalg.norm(a[0] - b[0]) < 1e-12 * np.linalg.norm(a[0])
