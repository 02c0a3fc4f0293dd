"""MI355X-native CheFSI + AAR-electrostatics hot path of SPARC.

sparc_amd.gpu        ctypes mirror of libsparcgpu.so (the C ABI in include/sparc_gpu.h)
sparc_amd.workloads  grid / projector containers and synthetic operator workloads
sparc_amd.build      hipcc build of libsparcgpu.so for gfx950
"""
from .workloads import Grid, NonlocalSet  # noqa: F401

__all__ = ["Grid", "NonlocalSet"]
