"""tetemoko_amd -- MI355X-native per-patch Godunov update for tetemoko.

The compute path is libtetemoko_hip.so (hand-written HIP for gfx950) behind the
C ABI in include/tetemoko_hip.h.  This package holds the ctypes binding of that
ABI (capi), seeded scenario inputs (scenarios), the multi-level driver (amr), the
scenario server (serving), result formats (stations, bdry_plot) and, under host/,
the C++ mirror of the reference's plugin classes (LinElastHIP.H).
"""
from . import capi, scenarios  # noqa: F401

__all__ = ["capi", "scenarios"]
