"""MI355X-native SFINCS explicit time step (momentum, continuity, CFL minimum dt).

The compute path is libsfx.so (hand-written HIP kernels for gfx950 behind a C ABI,
include/sfx.h); this package holds its ctypes bindings, the host mirror of the
reference's sfincs_lib entry points, and the seeded workload generators.
"""
from ._lib import (Batch, Instance, Model, SfxError, default_params, last_error, lib)  # noqa: F401
