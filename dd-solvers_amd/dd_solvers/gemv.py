"""Batched dense GEMV op, same name, schema and checks as the reference's
`dd_solvers::gemv_strided_batched` (src/dd_solvers/dd_solvers/gemv.py:4-30,
src/dd_solvers/cpp/gemv.cpp:9-57), running the HIP kernels of csrc/gemv.hip."""

from __future__ import annotations

from typing import Optional

import torch

from . import _lib

__all__ = ["gemv_strided_batched"]


def _check(matrices: torch.Tensor, vectors: torch.Tensor) -> None:
    # gemv.cpp:13-18
    if not (matrices.dim() == 3 and vectors.dim() == 2 and matrices.size(0) == vectors.size(0)
            and matrices.size(1) == matrices.size(2) and matrices.size(1) == vectors.size(1)):
        raise RuntimeError(
            "Input tensors must have compatible dimensions: (n, k, k) and (n, k)")
    if matrices.numel() > 0 and not (matrices.stride(1) == 1 and matrices.stride(2) == matrices.size(1)):
        raise RuntimeError("Matrices must be in column-major format")
    k = matrices.size(1)
    if matrices.size(0) > 1 and matrices.numel() > 0 and matrices.stride(0) != k * k:
        # the kernels address matrix b at b*k*k
        raise RuntimeError("gemv_strided_batched: the matrices must follow each other (batch stride k*k)")


@torch.library.custom_op("dd_solvers::gemv_strided_batched", mutates_args=())
def gemv_strided_batched(matrices: torch.Tensor, vectors: torch.Tensor,
                         use_shared_memory: Optional[bool] = None) -> torch.Tensor:
    """out[b] = matrices[b] @ vectors[b] for column-major (n, k, k) matrices.

    Accumulates in the vector dtype. `use_shared_memory=None` picks the LDS-staged kernel
    for 15 < k <= 1024 (float32 vectors) or 5 < k <= 1024 (float64), as gemv.py:8-14.
    """
    if use_shared_memory is None:
        k = vectors.shape[1]
        if vectors.dtype is torch.float32:
            use_shared_memory = 15 < k <= 1024
        elif vectors.dtype is torch.float64:
            use_shared_memory = 5 < k <= 1024
        else:
            raise ValueError("Unsupported data type")
    _check(matrices, vectors)
    _lib.require_device(matrices, vectors)
    vectors = vectors.contiguous()
    if matrices.dtype not in (torch.float64, torch.float32, torch.float16, torch.bfloat16) or \
            vectors.dtype not in (torch.float64, torch.float32):
        raise RuntimeError(f"unsupported dtypes: matrices {matrices.dtype}, vectors {vectors.dtype}")
    out = torch.empty_like(vectors)
    n, k = vectors.shape
    with torch.cuda.device(vectors.device):
        _lib.call("dds_gemv_strided_batched", _lib.dtype_code(matrices.dtype),
                  _lib.dtype_code(vectors.dtype), matrices.data_ptr(), vectors.data_ptr(),
                  out.data_ptr(), n, k, int(bool(use_shared_memory)), _lib.stream())
    return out


@gemv_strided_batched.register_fake
def _gemv_strided_batched_fake(matrices, vectors, use_shared_memory=None):
    if matrices.ndim != 3 or vectors.ndim != 2:
        raise RuntimeError("matrices must be 3D and vectors must be 2D")
    if matrices.shape[0] != vectors.shape[0]:
        raise RuntimeError("matrices and vectors must have the same batch size")
    return torch.empty_like(vectors)
