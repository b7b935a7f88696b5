"""MI355X-native HSC (Heterogeneous Subspace Correction) linear solver.

The product is the HIP library ``libhsc_mi355x.so`` behind the C ABI in
include/hsc.h; this package holds its sources (csrc/), the build recipe
(build.py), the ctypes binding (_abi.py), the plugin-shaped host mirror
(plugin.py) and the synthetic workload generator (systems.py).
"""
from . import _abi  # noqa: F401

__all__ = ["_abi", "plugin", "systems", "build"]
