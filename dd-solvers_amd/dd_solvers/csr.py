"""Sparse-format utilities (setup phase): CSR -> BSR, column sort, duplicate merge, and the
subdomain permutation helpers.

Same names, arguments and results as the reference's src/dd_solvers/dd_solvers/cusparse.py
(`csr_to_bsr` :74-75, `sort_csr` :25-71, `gather_csr` :78-158) and solvers.py
(`csr_to_torch` :116-122, `permute_csr_tensor` :125-167, `inverse_permutation` :306-309,
`asm_permutation` :541-551). They run as PyTorch index operations on the tensors' device;
none of them is on the solve path.
"""

from __future__ import annotations

import torch

from . import _lib

__all__ = [
    "csr_to_bsr",
    "sort_csr",
    "gather_csr",
    "csr_to_torch",
    "permute_csr_tensor",
    "inverse_permutation",
    "asm_permutation",
    "csr_rows",
    "segment_sum",
]


def segment_sum(vals: torch.Tensor, lengths: torch.Tensor,
                compensated: bool = False) -> torch.Tensor:
    """Sums of the consecutive runs of `vals` with the given lengths, each run added in
    index order in vals' dtype (float64 or float32); `compensated` carries the rounding error
    of each addition (Neumaier), for long runs whose terms cancel. On the GPU this is
    dds_segment_sum:
    torch.segment_reduce on ROCm returned wrong sums for 2^24 of the segments of a
    21-million-segment input (the merge of R_C A for the reference's p = 3 S9 / S9 / S9
    runs), so no setup step relies on it there."""
    n = int(lengths.shape[0])
    _lib.require_device(vals)
    ptr = torch.zeros(n + 1, dtype=torch.int64, device=vals.device)
    ptr[1:] = torch.cumsum(lengths.to(torch.int64), 0)
    vals = vals.contiguous()
    out = torch.empty(n, dtype=vals.dtype, device=vals.device)
    with torch.cuda.device(vals.device):
        _lib.call("dds_segment_sum", _lib.dtype_code(vals.dtype), int(compensated), n,
                  ptr.data_ptr(), vals.data_ptr(), out.data_ptr(), _lib.stream())
    return out


def csr_rows(crow: torch.Tensor) -> torch.Tensor:
    """Row index of every stored entry (int64)."""
    n = crow.shape[0] - 1
    counts = (crow[1:] - crow[:-1]).to(torch.int64)
    return torch.repeat_interleave(torch.arange(n, device=crow.device), counts)


def csr_to_bsr(matrix: torch.Tensor, block_size: int) -> torch.Tensor:
    """CSR -> BSR with block_size x block_size blocks stored column-major inside each block.

    Keeps every block that holds a stored entry (explicit zeros included), block columns
    ascending in each block row, crow/col int32, values strides (b*b, 1, b) — the layout of
    cusparse<S|D>csr2gebsr with CUSPARSE_DIRECTION_COLUMN (cusparse.cpp:22-107).
    """
    b = int(block_size)
    if b < 1:
        raise ValueError("block_size must be positive")
    crow, col, val = matrix.crow_indices(), matrix.col_indices(), matrix.values()
    _lib.require_device(val)
    nrows, ncols = matrix.shape
    dev = val.device
    nbr, nbc = -(-nrows // b), -(-ncols // b)
    row = csr_rows(crow)
    col64 = col.to(torch.int64)
    key = (row // b) * nbc + col64 // b
    ukey, inv = torch.unique(key, sorted=True, return_inverse=True)
    nnzb = ukey.shape[0]
    ubr, ubc = ukey // nbc, ukey % nbc
    bcrow = torch.zeros(nbr + 1, dtype=torch.int64, device=dev)
    bcrow[1:] = torch.cumsum(torch.bincount(ubr, minlength=nbr), 0)
    flat = torch.zeros(nnzb * b * b, dtype=val.dtype, device=dev)
    # column-major inside the block: (r, c) at c*b + r
    flat.index_put_((inv * (b * b) + (col64 % b) * b + row % b,), val, accumulate=True)
    values = flat.view(nnzb, b, b).transpose(1, 2)
    return torch.sparse_bsr_tensor(bcrow.to(torch.int32), ubc.to(torch.int32), values,
                                   size=(nbr * b, nbc * b))


def sort_csr(matrix: torch.Tensor, num_parts: int | None = None) -> torch.Tensor:
    """Sort the column indices inside every row (values follow); int32 indices out.

    `num_parts` splits the work into row chunks as the reference does to bound memory
    (cusparse.py:25-71); the result does not depend on it. The input is left intact.
    """
    rows, cols = matrix.shape
    _lib.require_device(matrix.values())
    crow = matrix.crow_indices().to(torch.int64)
    col = matrix.col_indices().to(torch.int64)
    val = matrix.values()
    if num_parts is None:
        num_parts = 1
    if not 1 <= num_parts <= max(rows, 1):
        raise ValueError(f"num_parts must be in [1, {rows}]")
    out_col = torch.empty_like(col)
    out_val = torch.empty_like(val)
    start = 0
    for part in range(num_parts):
        nr = rows // num_parts + (1 if part < rows % num_parts else 0)
        s, e = int(crow[start]), int(crow[start + nr])
        r = csr_rows(crow[start:start + nr + 1] - s)
        key = r * cols + col[s:e]
        order = torch.sort(key, stable=True).indices
        out_col[s:e] = col[s:e][order]
        out_val[s:e] = val[s:e][order]
        start += nr
    return torch.sparse_csr_tensor(crow.to(torch.int32), out_col.to(torch.int32), out_val,
                                   size=matrix.shape)


def gather_csr(matrix: torch.Tensor) -> torch.Tensor:
    """Merge runs of equal consecutive column indices in each row, summing their values
    (gather_csr_{count,sum}_kernel, cusparse.py:78-158). Index dtypes are preserved."""
    crow, col, val = matrix.crow_indices(), matrix.col_indices(), matrix.values()
    _lib.require_device(val)
    row = csr_rows(crow)
    nnz = col.shape[0]
    dev = val.device
    if nnz == 0:
        return torch.sparse_csr_tensor(crow.clone(), col.clone(), val.clone(), size=matrix.shape)
    new = torch.ones(nnz, dtype=torch.bool, device=dev)
    new[1:] = (col[1:] != col[:-1]) | (row[1:] != row[:-1])
    seg = torch.cumsum(new.to(torch.int64), 0) - 1
    nseg = int(seg[-1]) + 1
    lengths = torch.bincount(seg, minlength=nseg)
    sums = segment_sum(val, lengths)
    out_col = col[new]
    out_row = row[new]
    out_crow = torch.zeros_like(crow)
    out_crow[1:] = torch.cumsum(torch.bincount(out_row, minlength=crow.shape[0] - 1), 0).to(crow.dtype)
    return torch.sparse_csr_tensor(out_crow, out_col, sums, size=matrix.shape)


def csr_to_torch(csr) -> torch.Tensor:
    """scipy.sparse CSR -> torch sparse CSR (CPU)."""
    return torch.sparse_csr_tensor(torch.as_tensor(csr.indptr), torch.as_tensor(csr.indices),
                                   torch.as_tensor(csr.data), csr.shape)


def permute_csr_tensor(mat: torch.Tensor, perm: torch.Tensor) -> torch.Tensor:
    """CSR matrix Ap with Ap[perm][:, perm] == mat: entry (i, j) moves to (perm[i], perm[j]).
    Columns come back sorted, indices int32."""
    crow, col, val = mat.crow_indices(), mat.col_indices(), mat.values()
    _lib.require_device(val)
    perm = perm.to(device=val.device, dtype=torch.int64)
    n = mat.shape[0]
    new_row = perm[csr_rows(crow)]
    new_col = perm[col.to(torch.int64)]
    order = torch.sort(new_row * mat.shape[1] + new_col, stable=True).indices
    out_crow = torch.zeros(n + 1, dtype=torch.int64, device=val.device)
    out_crow[1:] = torch.cumsum(torch.bincount(new_row, minlength=n), 0)
    return torch.sparse_csr_tensor(out_crow.to(torch.int32), new_col[order].to(torch.int32),
                                   val[order], size=mat.shape)


def inverse_permutation(perm: torch.Tensor) -> torch.Tensor:
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(perm.shape[0], device=perm.device, dtype=perm.dtype)
    return inv


def asm_permutation(fine_to_solvers: torch.Tensor, solvers_to_coarse: torch.Tensor,
                    fine_to_dofs: torch.Tensor) -> torch.Tensor:
    """DoF order grouped by coarse cell, then by solver, then by fine element.

    Elements are ranked by the position of their solver in the (stable) coarse-cell order
    of the solvers; ties keep the element order. The element DoFs are listed in that
    order — the numbering of solvers.py:541-551.
    """
    dev = fine_to_dofs.device
    fine_to_solvers = fine_to_solvers.to(dev)
    solvers_to_coarse = solvers_to_coarse.to(dev)
    solver_order = torch.argsort(solvers_to_coarse, stable=True)
    rank = torch.empty_like(solver_order)
    rank[solver_order] = torch.arange(solver_order.shape[0], device=dev)
    element_order = torch.argsort(rank[fine_to_solvers.to(torch.int64)], stable=True)
    return fine_to_dofs[element_order].flatten()
