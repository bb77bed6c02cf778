"""The C-ABI library loads and exports every symbol include/lofar_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from lofar_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lofar_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_]\w*\s*\*?\s*([A-Za-z_]\w*)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_reference_dropins():
    names = declared_functions()
    for n in ("wstack", "phase_correction", "test", "lofar_fft_planes", "lofar_image_dev",
              "lofar_degrid_dev"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_host_only_entry_points():
    L = _lib.lib()
    assert L.test(41) == 42          # w-stacking_omp.c:241-247
    assert L.lofar_version() >= 100
    assert L.lofar_error_string(_lib.ECOORD).decode().startswith("sample")
    assert L.lofar_last_error() == 0
    assert L.lofar_stage_name(3).decode() == "grid_tiles"
