"""MI355X-native supernodal sparse Cholesky / static LDL^T (symPACK 2D path).

The compute path is libsympack_amd.so (HIP kernels for gfx950 behind a C ABI,
include/sympack_amd.h); this package only mirrors the symPACKMatrix2D
interface and provides the benchmark matrix generators.
"""
from . import matrices  # noqa: F401
from ._lib import NotHPDError, SymPACKError, load  # noqa: F401
from .matrix import (DistSparseMatrix, RelaxationParameters, compute_ordering,  # noqa: F401
                     symPACKMatrix2D, symPACKOptions)

__all__ = ["DistSparseMatrix", "RelaxationParameters", "symPACKMatrix2D", "symPACKOptions",
           "NotHPDError", "SymPACKError", "load", "matrices"]
