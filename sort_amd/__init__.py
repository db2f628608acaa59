"""sort_amd — MI355X-native ray–scene intersection and path tracing for SORT.

The compute path is libsortgpu.so (HIP kernels for gfx950 behind the C ABI in
include/sortgpu.h); this package holds its Python host side: the accelerator mirror
(``accel``), synthetic workload generators (``scenes``) and the render entry points.
"""
from ._lib import LIB_PATH, SortGpuError, header_symbols, lib  # noqa: F401
from .accel import MISS, Accelerator, Context, Obvh, Qbvh, hit_prims  # noqa: F401
