"""qlens_amd: MI355X-native pixellated-source log-evidence (QLens's inversion path).

The compute path is libqlx.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/qlx.h); this package holds the Python host mirror (qlx.PixelInversion), the build script
and the seeded workload definitions.
"""
from . import workloads  # noqa: F401
