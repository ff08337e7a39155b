"""MI355X-native Classic MPM DoublePass timestep (host package).

The timestep runs in hand-written gfx950 HIP kernels behind the C ABI of
include/mpm.h (libmpm_amd.so, built by ``python -m code_mpm_amd.build``);
this package is the Python host API over it.
"""
from . import capi, scenes  # noqa: F401
from .engine import Engine  # noqa: F401
