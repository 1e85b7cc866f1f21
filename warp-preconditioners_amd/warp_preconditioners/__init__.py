"""MI355X-native warp-preconditioners: FSAI, block Jacobi / ILU(0) and factored
biharmonic operators for CG/CR, on hand-written gfx950 HIP kernels.

Drop-in for the reference package's public surface (warp_preconditioners/__init__.py):
the same class names, constructor arguments, ``matvec(x, y, z, alpha, beta)``
semantics and error messages, with torch device tensors in place of Warp arrays
and ``warp_preconditioners.sparse.BsrMatrix`` in place of ``warp.sparse.BsrMatrix``.
"""

from .biharmonic import BiharmonicSystem, RepeatedFSAI
from .block_fsai import BlockFSAI, BsrBlockFSAI
from .block_ilu import BlockILU0, BlockJacobi
from .fsai import FSAI
from .mixed import MatchingSchur, MixedHarmonicSystem, ShiftedBlock
from .sparse_operator import SparseOperator
from .squared_laplacian import SquaredLaplacianOperator

__all__ = [
    "FSAI",
    "BiharmonicSystem",
    "RepeatedFSAI",
    "SparseOperator",
    "SquaredLaplacianOperator",
    "BlockJacobi",
    "BlockILU0",
    "MixedHarmonicSystem",
    "ShiftedBlock",
    "MatchingSchur",
    "BlockFSAI",
    "BsrBlockFSAI",
]
