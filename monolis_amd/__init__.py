"""monolis_amd -- MI355X (gfx950) implementation of monolis's Krylov hot path.

The compute path is libmonolis_hip.so (hand-written HIP kernels behind
monolis's flat C-ABI, see include/monolis_hip.h).  This package holds the
build recipe, the ctypes binding, a Python mirror of monolis's solver API, the
host-side BCSR container and the synthetic workload generators.
"""
from . import matrix, workloads  # noqa: F401
from ._lib import LIB_PATH, MonolisHipError, load  # noqa: F401

__all__ = ["matrix", "workloads", "load", "LIB_PATH", "MonolisHipError"]
