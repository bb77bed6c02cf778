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


def test_plan_create_rejects_sizes_before_touching_the_device():
    """An FP32C plan keeps sample indices in 32 bits: 2^32 samples return LOFAR_ESIZE, and an
    unsupported grid side too, from the argument checks alone (no device call, no plan)."""
    import ctypes as C
    L = _lib.lib()
    fake = C.c_void_p(64)  # never dereferenced: the checks fail first
    plan = C.c_void_p(12345)
    args = (2, 1 << 32, 1, 1, fake, fake, fake, fake, 1.0 / 64, 0.5, 7, 64, 64, None,
            C.byref(plan))
    assert L.lofar_plan_create(_lib.FP32C, *args) == _lib.ESIZE and not plan.value
    plan = C.c_void_p(12345)
    assert L.lofar_plan_create(_lib.FP64, 2, 10, 1, 1, fake, fake, fake, fake, 1.0 / 64, 0.5,
                               7, 64, 8193 * 2, None, C.byref(plan)) == _lib.ESIZE
    assert not plan.value
