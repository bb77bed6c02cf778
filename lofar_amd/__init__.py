"""lofar_amd: MI355X-native w-stacking imaging hot path (gridding, per-plane FFT,
w-phase correction and plane sum, and its adjoint) behind the reference's C ABI.

The compute lives in liblofar_hip.so (HIP kernels for gfx950); this package is
its ctypes binding, a host-side mirror of the reference interface and a seeded
synthetic-observation generator.
"""
from ._lib import FP32C, FP64, LIB_PATH, LibraryMissing, LofarError, lib  # noqa: F401

__all__ = ["FP32C", "FP64", "LIB_PATH", "LibraryMissing", "LofarError", "lib"]
