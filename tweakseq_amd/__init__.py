"""tweakseq_amd -- MI355X-native pairwise alignment engine behind tweakseq's
AlignmentTool boundary.

The product is libtweakalign.so (HIP kernels for gfx950 + C ABI,
include/tweakalign.h) and the `tweakalign` CLI beside it.  This package only
binds that library through ctypes for tests, benchmarks and Python callers;
there is no Python or CPU fallback of the alignment itself: without the
built library every entry point raises.
"""
from .engine import (  # noqa: F401
    Engine, Batch, Result, Workload, lib, encode, matrix, LOCAL, GLOBAL, PROTEIN, DNA,
    LANES_AUTO, LANES_I16, LANES_I32, LANES_LONG, F_I32, F_RERUN, F_LONG, F_EMPTY, F_WALK_FAILED, E_WALK,
)
