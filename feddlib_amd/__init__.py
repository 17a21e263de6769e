"""feddlib_amd -- MI355X-native element assembly for FEDDLib's FE/AssembleFE path.

The compute lives in libfeddgpu.so (hand-written HIP for gfx950 behind the C ABI
in include/feddgpu.h); this package holds its sources (csrc/), the build step
and a thin ctypes binding used by the tests and the benchmark.
"""
from . import native  # noqa: F401
