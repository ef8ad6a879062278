"""Oracle restatements of the reference's native ops (numpy, CPU).

Pins (tests/test_oracle.py):
  gemv_strided_batched  dense numpy matmul on the shapes of src/dd_solvers/tests/test_gemv.py
  bsrmv                 BSR fixture of src/dd_solvers/tests/test_bsr.py:9-138 vs dense
  csr_to_bsr            known answer of test_bsr.py:196-223 (crow, col, blocks, strides)
  sort_csr / gather_csr known answers of src/dd_solvers/tests/test_cusparse.py:8-78
  spd_solve             known answers of src/dd_solvers/tests/test_cudss.py:15-84, and
                        scipy.linalg.cho_solve on random SPD matrices
  local_csr_batch       scipy's slicing of the diagonal blocks A[ik:(i+1)k, ik:(i+1)k]
                        (tests/test_sparse_coarse.py), with an entry planted in the
                        wrong block caught
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

__all__ = [
    "round_bf16",
    "quantize",
    "gemv_strided_batched",
    "bsrmv",
    "csr_to_bsr",
    "sort_csr",
    "gather_csr",
    "cholesky_lower",
    "spd_solve",
    "local_csr_batch",
]


def round_bf16(x) -> np.ndarray:
    """Round to bfloat16 (round to nearest even on the float32 bits); returns float32."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32)
    nan = np.isnan(f)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def quantize(x, storage: str) -> np.ndarray:
    """Values of `x` after a round trip through the storage type, as float64."""
    x = np.asarray(x, dtype=np.float64)
    if storage in ("float64", "f64"):
        return x
    if storage in ("float32", "f32"):
        return x.astype(np.float32).astype(np.float64)
    if storage in ("float16", "f16"):
        return x.astype(np.float16).astype(np.float64)
    if storage in ("bfloat16", "bf16"):
        return round_bf16(x.astype(np.float32)).astype(np.float64)
    raise ValueError(storage)


def gemv_strided_batched(mat: np.ndarray, vec: np.ndarray, acc_dtype=None) -> np.ndarray:
    """y[b] = mat[b] @ vec[b], accumulated over columns c = 0..k-1 in the vector dtype.

    Follows src/dd_solvers/cpp/gemv_kernels.cu:31-57 (one output row per thread,
    `acc += upcast(m) * v` for ascending c, accumulator of the vector type) and the shape
    checks of src/dd_solvers/cpp/gemv.cpp:13-18. `mat` holds already-upcast values.
    """
    mat = np.asarray(mat)
    vec = np.asarray(vec)
    n, k = vec.shape
    if mat.shape != (n, k, k):
        raise RuntimeError("Input tensors must have compatible dimensions: (n, k, k) and (n, k)")
    dt = np.dtype(acc_dtype or vec.dtype)
    m = mat.astype(dt)
    v = vec.astype(dt)
    acc = np.zeros((n, k), dtype=dt)
    for c in range(k):
        acc += m[:, :, c] * v[:, c:c + 1]
    return acc


def bsrmv(crow, col, values, x) -> np.ndarray:
    """y = A x for BSR with values[block, r, c]: blocks of a row in stored order, then c.

    Follows the RPT-I kernel of src/dd_solvers/dd_solvers/bsr.py:46-66 (one scalar row per
    thread; the accumulator starts at a float64 literal, bsr.py:59).
    """
    crow = np.asarray(crow, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    values = np.asarray(values)
    b = values.shape[1]
    nbr = crow.shape[0] - 1
    x = np.asarray(x, dtype=np.float64)
    y = np.zeros(nbr * b, dtype=np.float64)
    for br in range(nbr):
        acc = np.zeros(b)
        for blk in range(crow[br], crow[br + 1]):
            off = col[blk] * b
            for c in range(b):
                acc += values[blk, :, c].astype(np.float64) * x[off + c]
        y[br * b:(br + 1) * b] = acc
    return y


def csr_to_bsr(indptr, indices, data, b: int, shape):
    """CSR -> BSR with b x b blocks; returns (crow, col, values[nnzb, b, b]).

    Follows the cusparse<S|D>csr2gebsr call of src/dd_solvers/cpp/cusparse.cpp:22-107:
    blocks holding at least one stored entry are kept (stored zeros count), block columns
    ascend inside a block row, and missing entries inside a kept block are zero. The
    reference stores each block column-major (cusparse.cpp:30, 81-83); the logical values
    returned here are values[blk, r, c] regardless of layout.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    data = np.asarray(data)
    nrows, ncols = shape
    nbr = -(-nrows // b)
    crow = [0]
    cols = []
    blocks = []
    for br in range(nbr):
        d = {}
        for r in range(br * b, min((br + 1) * b, nrows)):
            for p in range(indptr[r], indptr[r + 1]):
                c = indices[p]
                bc = c // b
                if bc not in d:
                    d[bc] = np.zeros((b, b), dtype=data.dtype)
                d[bc][r - br * b, c - bc * b] += data[p]
        for bc in sorted(d):
            cols.append(bc)
            blocks.append(d[bc])
        crow.append(len(cols))
    values = np.stack(blocks) if blocks else np.zeros((0, b, b), dtype=data.dtype)
    return np.array(crow), np.array(cols, dtype=np.int64), values


def sort_csr(indptr, indices, data):
    """Sort column indices inside each row, permuting the values alongside.

    Follows src/dd_solvers/dd_solvers/cusparse.py:25-71 with cusparseXcsrsort +
    cusparseGather (src/dd_solvers/cpp/cusparse.cpp:109-159); indices come back int32.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64).copy()
    data = np.asarray(data).copy()
    for r in range(indptr.shape[0] - 1):
        s, e = indptr[r], indptr[r + 1]
        o = np.argsort(indices[s:e], kind="stable")
        indices[s:e] = indices[s:e][o]
        data[s:e] = data[s:e][o]
    return indptr.astype(np.int32), indices.astype(np.int32), data


def gather_csr(indptr, indices, data):
    """Merge runs of equal consecutive column indices in each row by summing their values.

    Follows gather_csr_{count,sum}_kernel of src/dd_solvers/dd_solvers/cusparse.py:78-158.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices)
    data = np.asarray(data)
    out_ptr = [0]
    out_idx = []
    out_val = []
    for r in range(indptr.shape[0] - 1):
        prev = None
        for p in range(indptr[r], indptr[r + 1]):
            c = indices[p]
            if prev is not None and c == prev:
                out_val[-1] = out_val[-1] + data[p]
            else:
                out_idx.append(c)
                out_val.append(data[p])
                prev = c
        out_ptr.append(len(out_idx))
    return (np.array(out_ptr, dtype=indptr.dtype), np.array(out_idx, dtype=indices.dtype),
            np.array(out_val, dtype=data.dtype))


def cholesky_lower(A: np.ndarray) -> np.ndarray:
    """Lower Cholesky factor (LAPACK potrf through scipy)."""
    return sla.cholesky(np.asarray(A, dtype=np.float64), lower=True)


def spd_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x = A^{-1} b for SPD A by Cholesky factor and two triangular solves.

    Stands for the coarse-slot `CUDSS` solver (src/dd_solvers/dd_solvers/solvers.py:85-113,
    src/dd_solvers/cpp/cudss.cpp:64-105): an SPD sparse Cholesky whose result is the exact
    solve up to rounding.
    """
    c = sla.cho_factor(np.asarray(A, dtype=np.float64), lower=True)
    return sla.cho_solve(c, np.asarray(b, dtype=np.float64))


def local_csr_batch(indptr, indices, data, k: int):
    """The diagonal k x k blocks of a CSR matrix as the reference's CSRBatch fields
    (nrows, nnz, values, col_indices, row_offsets).

    Follows SchwarzOperator._construct_local_solvers_matrices_sparse
    (src/dd_solvers/dd_solvers/solvers.py:660-719) and its two kernels (:484-538): a row
    keeps its entries in the row's own subdomain, in the row's stored order, and stops at
    the first column of a later subdomain (the rows are sorted there); row_offsets has
    k + 1 entries per subdomain, counted from the subdomain's first entry.
    """
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    data = np.asarray(data)
    n = indptr.shape[0] - 1
    n_s = n // k
    count = np.zeros(n, dtype=np.int64)
    for row in range(n):
        for ind in range(indptr[row], indptr[row + 1]):
            col = indices[ind]
            if col // k == row // k:
                count[row] += 1
            elif col // k > row // k:
                break
    nnz = count.reshape(n_s, k).sum(axis=1)
    scan = np.concatenate([[0], np.cumsum(count)])
    values = np.empty(int(scan[-1]), dtype=data.dtype)
    cols = np.empty(int(scan[-1]), dtype=np.int32)
    row_offsets = np.empty(n + n_s, dtype=np.int32)
    for row in range(n):
        solver, solver_row = divmod(row, k)
        out = scan[row]
        first = scan[solver * k]
        if solver_row == 0:
            row_offsets[solver * (k + 1)] = out - first
        for ind in range(indptr[row], indptr[row + 1]):
            col = indices[ind]
            if col // k == row // k:
                cols[out] = col % k
                values[out] = data[ind]
                out += 1
            elif col // k > row // k:
                break
        row_offsets[solver * (k + 1) + solver_row + 1] = out - first
    return np.full(n_s, k, dtype=np.int32), nnz, values, cols, row_offsets
