"""seismic_amd: MI355X-native 2-D acoustic FWI propagator and inversion loops.

Hot path: hand-written gfx950 HIP kernels in libseis.so (C-ABI, include/seis.h)
driven from Python through ctypes; torch provides device tensors and streams.
"""
from .geometry import Geometry, ricker  # noqa: F401

__all__ = ["Geometry", "ricker"]
