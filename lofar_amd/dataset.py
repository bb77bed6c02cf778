"""The reference's on-disk inputs and outputs around the imaging hot path.

Readers and writers for the formats the reference's driver uses, so that a paramfile and a
binMS directory written for it run unchanged through the device path:

* ``read_paramfile``: the whitespace-separated ``key value`` parameter file
  (init.c:124-250), including its quirks (``DatapathN`` keys are only recognised in
  order once ``ndatasets`` > 1 has been read; ``verbose_level`` is parsed from the key,
  so it is always 0).
* ``read_meta``: ``meta.txt``, twelve numbers (init.c:275-296); ``metadata_calculation``
  drops the last 1000 measurements and sets the derived counts for one rank
  (init.c:315-349).
* ``read_data``: the binMS arrays ``ucoord.bin``, ``vcoord.bin``, ``wcoord.bin``
  (float64[N]), ``weights.bin`` (float32[N*P]) and ``visibilities_real.bin`` /
  ``visibilities_img.bin`` (float32[N*F*P]) (init.c:399-494); a short file leaves the
  rest zero, as the reference's calloc'd buffers do.
* ``sector_samples``: the one-rank sector selection of gridding.c:35-86,178-203 --
  samples with (int)vv == 0, gathered in reverse order.
* ``write_cube_dump`` / ``write_image``: the WRITE_DATA cube dumps (gridding.c:368-449,
  fourier_transform.c:76-161: float64 real and imaginary planes [w][v][u] in separate
  files) and the WRITE_IMAGE output (fourier_transform.c:177-219: float64 real and
  imaginary images [v][u]), under the dataset-index-prefixed names of init.c:70-112.
* ``run``: main.c's loop over datasets (read, grid, transform, phase-correct, write) on the
  device path.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List

import numpy as np

NSUB = 1000  # measurements dropped at the end (init.c:317)

_OUT_KEYS = ("outfile", "outfile1", "outfile2", "outfile3", "fftfile", "fftfile2", "fftfile3",
             "logfile", "extension", "timingfile")
_IN_KEYS = ("ufile", "vfile", "wfile", "weightsfile", "visrealfile", "visimgfile", "metafile")


@dataclasses.dataclass
class Params:
    """struct parameter, ip and op of allvars.h:51-128."""
    num_threads: int = 0
    ndatasets: int = 0
    datapath_multi: List[str] = dataclasses.field(default_factory=lambda: [""] * 10)
    grid_size_x: int = 0
    grid_size_y: int = 0
    num_w_planes: int = 0
    w_support: int = 0
    verbose_level: int = 0
    files: dict = dataclasses.field(default_factory=dict)    # in.*
    outparam: dict = dataclasses.field(default_factory=dict)  # outparam.*


def _atoi(s):
    """C atoi: leading whitespace, optional sign, digits; 0 if none."""
    s = s.strip()
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > i else 0


def read_paramfile(path) -> Params:
    """init.c:124-250: pairs of whitespace-separated tokens, in file order."""
    p = Params()
    with open(path, "r") as f:
        toks = f.read().split()
    i = 1
    for k in range(0, len(toks) - 1, 2):
        b1, b2 = toks[k], toks[k + 1]
        if b1 == "num_threads":
            p.num_threads = _atoi(b2)
        if b1 == "Datapath1":
            p.datapath_multi[0] = b2
            i += 1
        if b1 == "ndatasets":
            p.ndatasets = _atoi(b2)
        if b1 == "w_support":
            p.w_support = _atoi(b2)
        if b1 == "grid_size_x":
            p.grid_size_x = _atoi(b2)
        if b1 == "grid_size_y":
            p.grid_size_y = _atoi(b2)
        if b1 == "num_w_planes":
            p.num_w_planes = _atoi(b2)
        if b1 in _IN_KEYS:
            p.files[b1] = b2
        if b1 in _OUT_KEYS:
            p.outparam[b1] = b2
        if b1 == "verbose_level":
            p.verbose_level = _atoi(b1)  # the reference converts the key (init.c:240)
        if p.ndatasets > 1 and b1 == "Datapath%d" % i:
            p.datapath_multi[i - 1] = b2
            i += 1
    return p


@dataclasses.dataclass
class Meta:
    """struct meta of allvars.h:78-95."""
    Nmeasures: int
    Nvis: int
    freq_per_chan: int
    polarisations: int
    Ntimes: int
    dt: float
    thours: float
    baselines: int
    uvmin: float
    uvmax: float
    wmin: float
    wmax: float
    Nweights: int = 0
    startrow: int = 0


def read_meta(path) -> Meta:
    """init.c:282-296: %ld %ld %ld %ld %ld %lf %lf %ld %lf %lf %lf %lf."""
    with open(path, "r") as f:
        t = f.read().split()
    if len(t) < 12:
        raise ValueError("meta file %s holds %d of 12 numbers" % (path, len(t)))
    return Meta(_atoi(t[0]), _atoi(t[1]), _atoi(t[2]), _atoi(t[3]), _atoi(t[4]), float(t[5]),
                float(t[6]), _atoi(t[7]), float(t[8]), float(t[9]), float(t[10]), float(t[11]))


def metadata_calculation(meta: Meta) -> Meta:
    """init.c:315-349 for one rank: drop the last 1000 measurements, derive the counts."""
    m = dataclasses.replace(meta)
    m.Nmeasures = meta.Nmeasures - NSUB
    m.Nvis = m.Nmeasures * m.freq_per_chan * m.polarisations
    m.Nweights = m.Nmeasures * m.polarisations
    m.startrow = 0
    return m


def _read(path, dtype, count, offset_items):
    a = np.zeros(count, dtype)
    with open(path, "rb") as f:
        f.seek(offset_items * np.dtype(dtype).itemsize)
        got = np.fromfile(f, dtype=dtype, count=count)
    a[:len(got)] = got
    if len(got) != count:
        print("an error occurred while reading file %s" % path)
    return a


def read_data(datapath, params: Params, meta: Meta) -> dict:
    """init.c:399-494: the six binMS arrays of one dataset (after metadata_calculation)."""
    fp = meta.freq_per_chan * meta.polarisations
    n = meta.Nmeasures
    j = lambda key: os.path.join(datapath, params.files[key])  # noqa: E731
    return dict(
        uu=_read(j("ufile"), np.float64, n, meta.startrow),
        vv=_read(j("vfile"), np.float64, n, meta.startrow),
        ww=_read(j("wfile"), np.float64, n, meta.startrow),
        weight=_read(j("weightsfile"), np.float32, meta.Nweights,
                     meta.startrow * meta.polarisations),
        vis_real=_read(j("visrealfile"), np.float32, meta.Nvis, meta.startrow * fp),
        vis_img=_read(j("visimgfile"), np.float32, meta.Nvis, meta.startrow * fp),
    )


def sector_samples(data: dict, meta: Meta) -> dict:
    """gridding.c:35-86,178-203 with one rank (nsectors = 1): sector 0 holds the samples
    with (int)(vv*1) == 0, and its arrays are gathered from the last such sample to the
    first; vv is shifted by isector*shift = 0.  (A sample with (int)vv == 1 is counted into
    sector 1, which a single rank never grids.)"""
    vv = data["vv"]
    binphi = np.trunc(vv).astype(np.int64)
    idx = np.nonzero(binphi == 0)[0][::-1]
    F, P = meta.freq_per_chan, meta.polarisations
    vidx = (idx[:, None] * (F * P) + np.arange(F * P)[None, :]).reshape(-1)
    widx = (idx[:, None] * P + np.arange(P)[None, :]).reshape(-1)
    return dict(uu=np.ascontiguousarray(data["uu"][idx]), vv=np.ascontiguousarray(vv[idx]),
                ww=np.ascontiguousarray(data["ww"][idx]),
                weight=np.ascontiguousarray(data["weight"][widx]),
                vis_real=np.ascontiguousarray(data["vis_real"][vidx]),
                vis_img=np.ascontiguousarray(data["vis_img"][vidx]))


def out_name(params: Params, index: int, key: str, outdir="."):
    """init.c:70-112: output names are the dataset index followed by the parameter's name."""
    return os.path.join(outdir, "%d%s" % (index, params.outparam.get(key, key)))


def write_cube_dump(cube, real_path, imag_path):
    """WRITE_DATA layout (gridding.c:368-449, fourier_transform.c:76-161): float64 real and
    imaginary parts, [w][v][u], in two files."""
    c = np.asarray(cube)
    np.ascontiguousarray(c.real, dtype=np.float64).tofile(real_path)
    np.ascontiguousarray(c.imag, dtype=np.float64).tofile(imag_path)


def write_image(image_real, image_imag, real_path, imag_path):
    """WRITE_IMAGE layout (fourier_transform.c:177-219): float64 [v][u] real and imaginary."""
    np.ascontiguousarray(image_real, dtype=np.float64).tofile(real_path)
    np.ascontiguousarray(image_imag, dtype=np.float64).tofile(imag_path)


def write_dataset(datapath, meta: Meta, data: dict, params: Params = None):
    """Writes meta.txt and the six binMS arrays in the reference's formats (the inverse of
    read_meta/read_data; ``meta.Nmeasures`` is the count stored, before the 1000-row drop)."""
    os.makedirs(datapath, exist_ok=True)
    files = dict(ufile="ucoord.bin", vfile="vcoord.bin", wfile="wcoord.bin",
                 weightsfile="weights.bin", visrealfile="visibilities_real.bin",
                 visimgfile="visibilities_img.bin", metafile="meta.txt")
    if params is not None:
        files.update(params.files)
    with open(os.path.join(datapath, files["metafile"]), "w") as f:
        f.write("%d\n%d\n%d\n%d\n%d\n%.17g\n%.17g\n%d\n%.17g\n%.17g\n%.17g\n%.17g\n" % (
            meta.Nmeasures, meta.Nvis, meta.freq_per_chan, meta.polarisations, meta.Ntimes,
            meta.dt, meta.thours, meta.baselines, meta.uvmin, meta.uvmax, meta.wmin, meta.wmax))
    for key, arr, dt in (("ufile", "uu", np.float64), ("vfile", "vv", np.float64),
                         ("wfile", "ww", np.float64), ("weightsfile", "weight", np.float32),
                         ("visrealfile", "vis_real", np.float32),
                         ("visimgfile", "vis_img", np.float32)):
        np.ascontiguousarray(data[arr], dtype=dt).tofile(os.path.join(datapath, files[key]))
    return files


def run(paramfile, precision="fp64", outdir=".", write_data=False, write_image_files=True,
        device=None):
    """main.c's loop over the datasets of a paramfile on the device path: read
    (init.c), select the sector and grid (gridding.c), transform every plane
    (fourier_transform.c:7-74), phase-correct and sum (:164-173), and write the image
    files (WRITE_IMAGE) and, with write_data, the gridded and transformed cubes
    (WRITE_DATA).  Returns a list of (image_real, image_imag) numpy arrays."""
    import torch

    from . import imager

    dev = torch.device(device or "cuda:0")
    params = read_paramfile(paramfile)
    if params.num_threads == 0:
        raise ValueError("num_threads is 0 in %s (main.c:61-65)" % paramfile)
    images = []
    nx, ny, nw = params.grid_size_x, params.grid_size_y, params.num_w_planes
    for index in range(params.ndatasets):
        datapath = params.datapath_multi[index]
        meta = metadata_calculation(
            read_meta(os.path.join(datapath, params.files.get("metafile", "meta.txt"))))
        data = read_data(datapath, params, meta)
        sec = sector_samples(data, meta)
        res = imager.resolution_from_meta(meta.uvmin, meta.uvmax)
        t = {k: torch.from_numpy(v).to(dev) for k, v in sec.items()}
        vis = (t["uu"], t["vv"], t["ww"], t["vis_real"], t["vis_img"], t["weight"])
        kw = dict(freq_per_chan=meta.freq_per_chan, polarizations=meta.polarisations)
        if write_data:
            cube = imager.grid(*vis, nw, params.w_support, nx, ny, precision=precision, **kw)
            write_cube_dump(cube.cpu().numpy(), out_name(params, index, "outfile2", outdir),
                            out_name(params, index, "outfile3", outdir))
            cube = imager.fft(cube, out=cube)
            write_cube_dump(cube.cpu().numpy(), out_name(params, index, "fftfile2", outdir),
                            out_name(params, index, "fftfile3", outdir))
            re, im = imager.phase(cube, res, meta.wmin, meta.wmax)
        else:
            re, im = imager.image(*vis, nw, params.w_support, nx, ny, res, meta.wmin, meta.wmax,
                                  precision=precision, **kw)
        re, im = re.cpu().numpy(), im.cpu().numpy()
        if write_image_files:
            # with WRITE_DATA the fft files hold the transformed cube; the reference then
            # overwrites them with the image (fourier_transform.c:181-182)
            write_image(re, im, out_name(params, index, "fftfile2", outdir),
                        out_name(params, index, "fftfile3", outdir))
        images.append((re, im))
    return images


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run the reference's imaging driver on the device path")
    ap.add_argument("paramfile")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"])
    ap.add_argument("--outdir", default=".")
    ap.add_argument("--write-data", action="store_true",
                    help="also write the gridded and transformed cubes (WRITE_DATA)")
    a = ap.parse_args(argv)
    run(a.paramfile, a.precision, a.outdir, a.write_data)


if __name__ == "__main__":
    main()
