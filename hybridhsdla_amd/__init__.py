"""hybridhsdla_amd - MI355X-native FLAPW muffin-tin H/S generation.

The hot path (A/B matching coefficients and the per-atom contraction into the
Hermitian complex128 H and S) runs as hand-written HIP kernels for gfx950
behind the C ABI in include/hsgen.h (libhsgen.so).  This package holds the
HIP sources (csrc/), the build script, the ctypes binding, the system
description and the seeded synthetic workloads.
"""
from .system import KPoint, System, gvectors, kmax_for_ng, make_config, make_system, tiny_system

__all__ = ["KPoint", "System", "gvectors", "kmax_for_ng", "make_config", "make_system",
           "tiny_system", "Context", "HsgenError"]


def __getattr__(name):
    # the GPU binding is imported lazily so that CPU-only tooling can use the
    # system description without loading libhsgen.so
    if name in ("Context", "HsgenError", "SystemHandle"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
