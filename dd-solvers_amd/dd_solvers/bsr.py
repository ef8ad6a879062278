"""`FastBSR`: BSR matrix-vector product on HIP kernels.

Keeps the reference's class name, constructor and `@` operator
(src/dd_solvers/dd_solvers/bsr.py:15-32). Backends:
  "row_per_thread_irrespective"  the default (bsr.py:17). For b in {2, 4, 8} it runs the
                                 segment kernel (lanes stream the block row's contiguous
                                 values); for b in {3, 5, 6, 7} one lane per block with a
                                 butterfly over a block row's lanes; other b use one lane
                                 per scalar row, the RPT-I order (bsr.py:46-66).
  "torch"                        torch's own BSR @ dense (library comparator).
"""

from __future__ import annotations

import torch

from . import _lib

__all__ = ["FastBSR"]


class FastBSR:
    def __init__(self, matrix: torch.Tensor, backend: str = "row_per_thread_irrespective"):
        if matrix.layout is not torch.sparse_bsr:
            raise ValueError("FastBSR expects a torch.sparse_bsr tensor")
        _lib.require_device(matrix.values())
        if backend not in self.matmul_backends:
            raise ValueError(f"unknown backend {backend!r}; one of {list(self.matmul_backends)}")
        self.backend = backend
        self.matrix = matrix
        self.shape = matrix.shape
        vals = matrix.values()
        if vals.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"unsupported BSR dtype {vals.dtype}")
        self.block_size = int(vals.shape[1])
        self.n_block_rows = int(matrix.crow_indices().shape[0] - 1)
        self.dtype = vals.dtype
        self.device = vals.device
        crow = matrix.crow_indices()
        self.crow = crow if crow.dtype in (torch.int32, torch.int64) else crow.to(torch.int64)
        self.crow = self.crow.contiguous()
        self.col = matrix.col_indices().to(torch.int32).contiguous()
        # Column-major storage inside each block: memory [blk][c][r].
        self.values_cm = vals.transpose(1, 2).contiguous()
        self.nnzb = int(self.col.shape[0])

    # -- the product --------------------------------------------------------------------
    def __matmul__(self, x: torch.Tensor) -> torch.Tensor:
        """A @ x. A vector shorter than the block-padded width is zero-padded and the
        result cut back to its length (BSR of an n x n matrix with b not dividing n)."""
        n = x.shape[0]
        if n < self.shape[1]:
            xp = torch.zeros(self.shape[1], dtype=x.dtype, device=x.device)
            xp[:n] = x
            x = xp
        if self.backend == "torch":
            return (self.matrix @ x)[:n]
        y = torch.empty(self.shape[0], dtype=self.dtype, device=self.device)
        with torch.cuda.device(self.device):
            self._enqueue_matvec(x.contiguous(), y, None, None)
        return y[:n]

    def _enqueue_matvec(self, x, y, partials, gate) -> int:
        """y = A x on the current stream; with `partials` also x . y per-block sums.
        Returns the number of partials written."""
        b = self.block_size
        with torch.cuda.device(self.device):
            return _lib.call_partials(
                "dds_bsrmv", _lib.dtype_code(self.dtype), _lib.dtype_code(self.crow.dtype),
                self.n_block_rows, b, self.crow.data_ptr(), self.col.data_ptr(),
                self.values_cm.data_ptr(), x.data_ptr(), y.data_ptr(), _lib.ptr(partials),
                _lib.PARTIALS_OUT, _lib.ptr(gate), _lib.stream())

    @property
    def bytes_per_matvec(self) -> int:
        """Bytes one product reads and writes: values, block columns, block-row pointers,
        x (once) and y."""
        e = self.values_cm.element_size()
        return (self.nnzb * self.block_size ** 2 * e + 4 * self.nnzb
                + self.crow.element_size() * (self.n_block_rows + 1) + 2 * e * self.shape[0])

    matmul_backends = {"row_per_thread_irrespective": None, "torch": None}
