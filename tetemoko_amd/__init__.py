"""tetemoko_amd -- MI355X-native per-patch Godunov update for tetemoko.

The compute path is libtetemoko_hip.so (hand-written HIP for gfx950) behind the
C ABI in include/tetemoko_hip.h.  This package holds the ctypes binding of that
ABI (capi), seeded scenario inputs (scenarios) and the host-side mirror of the
reference's plugin classes (plugin).
"""
from . import capi, scenarios  # noqa: F401

__all__ = ["capi", "scenarios"]
