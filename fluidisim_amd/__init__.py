"""Fluidisim δ-3D pressure projection on MI355X (gfx950).

The product is libfsp.so (HIP kernels + C ABI, include/fsp.h); this package
holds its ctypes bindings, a host-side mirror of the reference's Volume3
projection interface and the synthetic workloads.
"""
from . import _lib
from .volume import Domain3, Error, ProjectionError, Report, Volume3, generate

__all__ = ["Domain3", "Error", "ProjectionError", "Report", "Volume3", "generate", "_lib"]
