"""bullet3_amd -- MI355X-native Bullet3 GPU rigid-body step.

The product is the C-ABI library bullet3_amd/libb3hip.so (HIP kernels for
gfx950 + C++ host code, declared in include/b3hip.h).  This package holds the
Python-side plumbing used by tests and benchmarks: the ctypes binding (capi),
record dtypes (layout), seeded scene generators (scenes) and a mirror of the
reference demo set-up (world).
"""
from . import capi, layout, scenes  # noqa: F401

__all__ = ["capi", "layout", "scenes", "world"]
