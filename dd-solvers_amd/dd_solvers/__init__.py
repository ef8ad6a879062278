"""dd_solvers for MI355X: Additive/Hybrid Schwarz-preconditioned CG on hand-written HIP
kernels (gfx950), with the Python API of mgr0dzicki/dd-solvers
(src/dd_solvers/dd_solvers/__init__.py: `from .solvers import *`).

Importing the package loads libdd_solvers_hip.so and the op module _C (which registers the
reference's native op torch.ops.dd_solvers_gemv.gemv_strided_batched, as the reference's
`from . import _C` does) and fails if they are not built.
"""

from . import _lib  # noqa: F401  (fails loudly when the HIP library is missing)
from . import _C  # noqa: F401  (registers torch.ops.dd_solvers_gemv)
from .solvers import *  # noqa: F401,F403
from .solvers import __all__ as _solvers_all
from .bsr import FastBSR
from .csr import csr_to_bsr, gather_csr, sort_csr
from .gemv import gemv_strided_batched
from .sparse_coarse import BatchSparseCholesky, SparseCholesky

__all__ = [*_solvers_all, "FastBSR", "csr_to_bsr", "gather_csr", "sort_csr",
           "gemv_strided_batched", "SparseCholesky", "BatchSparseCholesky"]
