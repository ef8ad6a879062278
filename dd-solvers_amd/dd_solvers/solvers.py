"""Solvers: CG, Additive/Hybrid Schwarz preconditioners, local and coarse solvers.

Drop-in for src/dd_solvers/dd_solvers/solvers.py: the same classes, constructor arguments,
`setup` / `solve` / `destroy` contract, metadata keys and exceptions. The arithmetic runs in
the HIP kernels of csrc/ through the C ABI (_lib); PyTorch only provides device memory and
the current stream.

What changes underneath (the contract does not):
  * CG keeps every scalar and the norm test on the device (dds_pcg_state) and reads the
    state back every `check_every` iterations instead of every iteration; kernels enqueued
    after convergence are no-ops, so iterates, norms and counts are those of the
    reference's loop (solvers.py:228-297). Optionally the iteration block is captured in a
    hipGraph and replayed.
  * The coarse solver slot gets `DenseCholesky`, a hand-written dense Cholesky factor and
    triangular solve, instead of cuDSS.
  * Subdomains may be uneven and overlapping (keyword `decomposition=` of `setup`); the
    reference's equal, contiguous, non-overlapping case is the default.
"""

from __future__ import annotations

import gc
from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Any, NamedTuple, Tuple

import numpy as np
import torch

from . import _lib
from .bsr import FastBSR
from .csr import (csr_rows, csr_to_bsr, csr_to_torch, inverse_permutation, permute_csr_tensor,
                  segment_sum)
from .csr import asm_permutation  # noqa: F401  (re-exported)
from .gemv import gemv_strided_batched

__all__ = [
    "Metadata",
    "ConvergenceNotAchieved",
    "collect",
    "timeit",
    "SparseSolver",
    "DenseCholesky",
    "CG",
    "DenseBatchSolver",
    "CSRBatch",
    "SparseBatchSolver",
    "Inv",
    "LU",
    "Cholesky",
    "SchwarzOperator",
    "AdditiveSchwarz",
    "HybridSchwarz",
    "csr_to_torch",
    "permute_csr_tensor",
    "inverse_permutation",
    "asm_permutation",
]

Metadata = dict[str, Any]


class ConvergenceNotAchieved(Exception):
    """Raised by CG when the iteration limit is reached (solvers.py:28-37)."""

    def __init__(self, metadata: Metadata):
        self.metadata = metadata
        if "iterations" not in self.metadata:
            msg = "Convergence not achieved."
        else:
            msg = f"Convergence not achieved in {self.metadata['iterations']} iterations."
        super().__init__(msg)


def collect():
    """Collect garbage and return cached device memory (solvers.py:40-45)."""
    gc.collect()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()


def timeit(fun, warmup_iters=10, iters=10, repetitions=20):
    """Minimum over repetitions of the mean event-timed duration in ms (solvers.py:48-65)."""
    times = []
    for _ in range(repetitions):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        for _ in range(warmup_iters):
            fun()
        start.record()
        for _ in range(iters):
            fun()
        end.record()
        torch.cuda.synchronize()
        times.append(start.elapsed_time(end) / iters)
    return min(times)


class SparseSolver(ABC):
    @abstractmethod
    def setup(self, matrix: torch.Tensor, *args, **kwargs) -> Metadata | None:
        """Set the solver up for the given matrix (sparse CSR)."""

    @abstractmethod
    def solve(self, rhs: torch.Tensor) -> Tuple[torch.Tensor, Metadata]:
        pass

    @abstractmethod
    def destroy(self) -> None:
        pass


def _dense_from(matrix: torch.Tensor) -> torch.Tensor:
    if matrix.layout in (torch.sparse_csr, torch.sparse_coo, torch.sparse_bsr, torch.sparse_csc):
        return matrix.to_dense()
    return matrix


# =======================================================================================
# Coarse solver
# =======================================================================================


class DenseCholesky(SparseSolver):
    """Coarse-space slot: dense Cholesky factorisation and triangular solves in HIP.

    Replaces the reference's `CUDSS()` (solvers.py:85-113), an SPD sparse Cholesky of the
    coarse matrix. The coarse matrices of these solvers are small and dense-ish (n_c up
    to a few thousand), so the factor is stored dense (row-major, lower triangle) and
    computed by a blocked right-looking Cholesky (dds_chol_factor).

    mode "trsv"    each solve runs forward and backward substitution with L in one
                   workgroup (dds_chol_solve): the least traffic, but a sequential chain
                   of 2 * n / 64 dependent steps on one CU.
    mode "inverse" setup also forms A^-1 = L^-T L^-1 by triangular solves with the factor
                   (dds_chol_inverse) and each solve is one dense symmetric matvec spread
                   over all CUs (dds_dense_symv).
    mode "auto"    "inverse" from `inverse_threshold` unknowns up, else "trsv". On MI355X
                   the matvec wins from tiny sizes on (n = 512: a few us against ~0.3 ms
                   for the one-workgroup substitution), so the default threshold is 32.
    """

    def __init__(self, mode: str = "auto", inverse_threshold: int = 32):
        if mode not in ("auto", "trsv", "inverse"):
            raise ValueError(f"unknown mode {mode!r}")
        self.mode = mode
        self.inverse_threshold = inverse_threshold
        self.L = None
        self.Ainv = None

    def __str__(self):
        return "DenseCholesky"

    def setup(self, matrix: torch.Tensor, *args, **kwargs) -> None:
        dense = _dense_from(matrix)
        _lib.require_device(dense)
        if dense.dtype not in (torch.float32, torch.float64):
            raise ValueError(f"unsupported coarse dtype {dense.dtype}")
        n = dense.shape[0]
        if dense.shape != (n, n):
            raise ValueError("coarse matrix must be square")
        self.dtype = dense.dtype
        self.n = n
        self.L = dense.contiguous().clone()
        info = torch.zeros(1, dtype=torch.int32, device=dense.device)
        with torch.cuda.device(dense.device):
            _lib.call("dds_chol_factor", _lib.dtype_code(self.dtype), n, self.L.data_ptr(), n,
                      info.data_ptr(), _lib.stream())
            bad = int(info.item())
            if bad:
                raise np.linalg.LinAlgError(
                    f"coarse matrix is not positive definite (pivot {bad - 1})")
            mode = self.mode
            if mode == "auto":
                mode = "inverse" if n >= self.inverse_threshold else "trsv"
            if mode == "trsv" and n > 6144:
                mode = "inverse"
            self.active_mode = mode
            # the explicit inverse serves "inverse" single solves and every block solve
            self.Ainv = torch.empty_like(self.L)
            _lib.call("dds_chol_inverse", _lib.dtype_code(self.dtype), n, self.L.data_ptr(),
                      n, self.Ainv.data_ptr(), n, _lib.stream())

    def solve(self, rhs: torch.Tensor) -> Tuple[torch.Tensor, Metadata]:
        if rhs.dtype != self.dtype:
            raise AssertionError(f"rhs dtype {rhs.dtype} != coarse dtype {self.dtype}")
        _lib.require_device(rhs)
        rhs = rhs.contiguous()
        y = torch.empty_like(rhs)
        with torch.cuda.device(rhs.device):
            self._enqueue_solve(rhs, y, None)
        return y, {}

    def _enqueue_solve_multi(self, x, y, st) -> None:
        """Y = A^-1 X for 16 columns (FP64)."""
        if self.dtype != torch.float64:
            raise ValueError("block coarse solves run in float64")
        _lib.call("dds_dense_symm_m", self.n, self.Ainv.data_ptr(), x.data_ptr(), y.data_ptr(),
                  _lib.ptr(st), _lib.stream())

    def _enqueue_solve(self, x, y, gate) -> None:
        dt = _lib.dtype_code(self.dtype)
        if self.active_mode == "inverse":
            _lib.call("dds_dense_symv", dt, self.n, self.Ainv.data_ptr(), self.n, x.data_ptr(),
                      y.data_ptr(), _lib.ptr(gate), _lib.stream())
        else:
            _lib.call("dds_chol_solve", dt, self.n, self.L.data_ptr(), self.n, x.data_ptr(),
                      y.data_ptr(), _lib.ptr(gate), _lib.stream())

    @property
    def bytes_per_solve(self) -> int:
        e = 8 if self.dtype == torch.float64 else 4
        if self.active_mode == "inverse":
            return e * self.n * self.n + 2 * e * self.n
        return e * self.n * (self.n + 1) + 4 * e * self.n  # two sweeps of the triangle

    def destroy(self):
        self.L = None
        self.Ainv = None


# =======================================================================================
# Operators
# =======================================================================================


class _CSROperator:
    """y = A x for a torch sparse CSR matrix (dds_csrmv); used by CG(bsr_matmul=False)
    and for R0A in HybridSchwarz."""

    def __init__(self, matrix: torch.Tensor, lanes: int | None = None):
        if matrix.layout is not torch.sparse_csr:
            matrix = matrix.to_sparse_csr()
        vals = matrix.values()
        _lib.require_device(vals)
        if vals.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"unsupported CSR dtype {vals.dtype}")
        self.shape = matrix.shape
        self.dtype = vals.dtype
        self.device = vals.device
        crow = matrix.crow_indices()
        self.crow = (crow if crow.dtype in (torch.int32, torch.int64) else crow.to(torch.int64)).contiguous()
        self.col = matrix.col_indices().to(torch.int32).contiguous()
        self.values = vals.contiguous()
        self.nnz = int(self.col.shape[0])
        avg = self.nnz / max(self.shape[0], 1)
        if lanes is None:
            # one wavefront's lanes per row up to 64 (128 entries on average); longer rows
            # (R_C A: few rows of hundreds of entries) get a workgroup of 256 per row
            cap = 256 if avg > 128 else 64
            lanes = 1
            while lanes < cap and lanes < avg:
                lanes *= 2
        self.lanes = int(lanes)

    def __matmul__(self, x):
        y = torch.empty(self.shape[0], dtype=self.dtype, device=self.device)
        with torch.cuda.device(self.device):
            self._enqueue_matvec(x.contiguous(), y, None, None)
        return y

    def _enqueue_matvec(self, x, y, partials, gate) -> int:
        if partials is not None and self.shape[0] != self.shape[1]:
            raise ValueError("the x . y epilogue needs a square matrix")
        return _lib.call_partials(
            "dds_csrmv", _lib.dtype_code(self.dtype), _lib.dtype_code(self.crow.dtype),
            self.shape[0], self.crow.data_ptr(), self.col.data_ptr(), self.values.data_ptr(),
            x.data_ptr(), y.data_ptr(), self.lanes, _lib.ptr(partials), _lib.PARTIALS_OUT,
            _lib.ptr(gate), _lib.stream())

    @property
    def bytes_per_matvec(self) -> int:
        e = self.values.element_size()
        return (self.nnz * (e + 4) + self.crow.element_size() * (self.shape[0] + 1)
                + e * (self.shape[0] + self.shape[1]))


class _SELLOperator:
    """y = A x in SELL-64 (dds_sellmv), built from a CSR matrix at setup: rows in slices
    of 64, each slice padded to its longest row and stored slot by slot, two slots
    interleaved per pair, so a wavefront reads a slot pair of its 64 rows with one
    coalesced 16-byte load per lane. CG(bsr_matmul=False) uses it for square matrices."""

    def __init__(self, matrix: torch.Tensor, lanes: int | None = None):
        if matrix.layout is not torch.sparse_csr:
            matrix = matrix.to_sparse_csr()
        vals = matrix.values()
        _lib.require_device(vals)
        if vals.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"unsupported CSR dtype {vals.dtype}")
        self.shape = matrix.shape
        self.dtype = vals.dtype
        self.device = dev = vals.device
        n = self.shape[0]
        crow = matrix.crow_indices().to(torch.int64)
        cnt = crow[1:] - crow[:-1]
        n_slices = -(-n // 64)
        padded = torch.zeros(n_slices * 64, dtype=torch.int64, device=dev)
        padded[:n] = cnt
        width = padded.view(n_slices, 64).max(dim=1).values
        slice_ptr = torch.zeros(n_slices + 1, dtype=torch.int64, device=dev)
        slice_ptr[1:] = torch.cumsum(64 * width, 0)
        total = int(slice_ptr[-1])
        slot = torch.arange(total, device=dev)
        q = torch.searchsorted(slice_ptr, slot, right=True) - 1
        slot_row = q * 64 + (slot - slice_ptr[q]) % 64
        col = slot_row.clamp(max=n - 1).to(torch.int32)
        val = torch.zeros(total, dtype=vals.dtype, device=dev)
        row = csr_rows(crow)
        t = torch.arange(row.shape[0], device=dev) - crow[row]
        dest = slice_ptr[row // 64] + 64 * t + row % 64
        col[dest] = matrix.col_indices().to(torch.int32)
        val[dest] = vals
        self.slice_ptr, self.col, self.values = slice_ptr, col, val
        self.nnz = int(vals.shape[0])
        self.stored = total
        self._compact_columns(width)
        self._pair_slots(width)

    def _pair_slots(self, width):
        """Reorder every stored array from slot order (entry (j, l) of a slice at 64 j + l)
        into slot pairs: (j, l) at 128 (j // 2) + 2 l + j % 2 for j < 2 (w // 2), an odd
        last slot unchanged, so a lane loads its two entries of a slot pair at once."""
        dev = self.device
        sp = self.slice_ptr
        q = torch.repeat_interleave(torch.arange(width.shape[0], device=dev), 64 * width)
        e = torch.arange(self.stored, device=dev) - sp[q]
        j, l = e // 64, e % 64
        paired = j < 2 * (width[q] // 2)
        new = torch.where(paired, 128 * (j // 2) + 2 * l + j % 2, e)

        def move(a, base):
            out = torch.empty_like(a)
            out[base + new] = a[base + e]
            return out

        self.values = move(self.values, sp[q])
        self.col = move(self.col, sp[q])
        cp = self.col_ptr[q]
        n16 = cp >= 0
        if self.col16.numel():
            c16 = torch.empty_like(self.col16)
            c16[cp[n16] + new[n16]] = self.col16[cp[n16] + e[n16]]
            self.col16 = c16
        if self.col32.numel():
            c32 = torch.empty_like(self.col32)
            o = -1 - cp[~n16]
            c32[o + new[~n16]] = self.col32[o + e[~n16]]
            self.col32 = c32

    def _compact_columns(self, width):
        """Per slot (64 entries) an int32 base and 16-bit offsets when the slot's columns
        span < 65536; slices with a wider slot keep int32 columns (dds_sellmv)."""
        dev = self.device
        n_slices = width.shape[0]
        C = self.col.view(-1, 64).to(torch.int64)
        base = C.min(dim=1).values
        span = C.max(dim=1).values - base
        slot_slice = torch.repeat_interleave(torch.arange(n_slices, device=dev), width)
        wide = torch.zeros(n_slices, dtype=torch.int64, device=dev).scatter_reduce_(
            0, slot_slice, (span > 65535).to(torch.int64), "amax").bool()
        entries = 64 * width
        n16 = torch.where(wide, 0, entries)
        n32 = torch.where(wide, entries, 0)
        off16 = torch.cumsum(n16, 0) - n16
        off32 = torch.cumsum(n32, 0) - n32
        self.col_ptr = torch.where(wide, -1 - off32, off16).contiguous()
        slot_wide = wide[slot_slice]
        d = (C[~slot_wide] - base[~slot_wide, None]).reshape(-1)
        self.col16 = torch.where(d >= 32768, d - 65536, d).to(torch.int16).contiguous()
        self.col32 = C[slot_wide].reshape(-1).to(torch.int32).contiguous()
        self.slot_base = torch.where(slot_wide, torch.zeros_like(base), base).to(torch.int32).contiguous()
        self.wide_fraction = float(wide.float().mean()) if n_slices else 0.0

    def __matmul__(self, x):
        y = torch.empty(self.shape[0], dtype=self.dtype, device=self.device)
        with torch.cuda.device(self.device):
            self._enqueue_matvec(x.contiguous(), y, None, None)
        return y

    def _enqueue_matvec(self, x, y, partials, gate) -> int:
        return _lib.call_partials(
            "dds_sellmv", _lib.dtype_code(self.dtype), self.shape[0], self.slice_ptr.data_ptr(),
            self.col_ptr.data_ptr(), _lib.ptr(self.col16), _lib.ptr(self.col32),
            self.slot_base.data_ptr(), self.values.data_ptr(), x.data_ptr(), y.data_ptr(),
            _lib.ptr(partials), _lib.PARTIALS_OUT, _lib.ptr(gate), _lib.stream())

    def _enqueue_matmat(self, X, Y, partials, st) -> int:
        """Y = A X for (n, 16) blocks; with `partials` also per-column X . Y sums."""
        if self.dtype != torch.float64:
            raise ValueError("block products run in float64")
        return _lib.call_partials(
            "dds_sellmm", self.shape[0], self.slice_ptr.data_ptr(), self.col.data_ptr(),
            self.values.data_ptr(), X.data_ptr(), Y.data_ptr(), _lib.ptr(partials),
            _lib.PARTIALS_OUT, _lib.ptr(st), _lib.stream())

    @property
    def bytes_stored(self) -> int:
        """Bytes one matvec streams from the stored format (values, compact columns, slot
        bases, slice pointers) plus x and y."""
        e = self.values.element_size()
        return (self.stored * e + 2 * self.col16.numel() + 4 * self.col32.numel()
                + 4 * self.slot_base.numel() + 16 * self.col_ptr.numel() + e * 2 * self.shape[0])

    @property
    def bytes_per_matvec(self) -> int:
        """CSR-equivalent (useful) bytes: values + column indices of the stored entries,
        row pointers, x and y; SELL padding is not credited."""
        e = self.values.element_size()
        return self.nnz * (e + 4) + 8 * (self.shape[0] + 1) + e * (self.shape[0] + self.shape[1])

    @property
    def bytes_physical(self) -> int:
        """Bytes the kernel itself moves at least once: every stored slot (padding
        included) of the values and of the compact columns (2 bytes per entry, 4 in wide
        slices), the per-slot column bases, the slice tables, x and y."""
        e = self.values.element_size()
        n_slices = self.slice_ptr.shape[0] - 1
        return (self.stored * e + 2 * self.col16.numel() + 4 * self.col32.numel()
                + 4 * self.slot_base.numel() + 16 * n_slices + e * (self.shape[0] + self.shape[1]))


class _TorchOperator:
    """Fallback operator for matrices the kernels do not take (dense, other layouts)."""

    def __init__(self, matrix):
        self.m = matrix
        self.shape = matrix.shape

    def __matmul__(self, x):
        return self.m @ x


# =======================================================================================
# CG
# =======================================================================================

_STATE_DTYPE = np.dtype([("atol", "f8"), ("rho", "f8"), ("rho_prev", "f8"), ("pq", "f8"),
                         ("alpha", "f8"), ("beta", "f8"), ("rr", "f8"), ("it", "i4"),
                         ("done", "i4"), ("maxiter", "i4"), ("pad", "i4")])
assert _STATE_DTYPE.itemsize == 72


def _state_tensor(atol: float, maxiter: int, device) -> torch.Tensor:
    s = np.zeros(1, dtype=_STATE_DTYPE)
    s["atol"] = atol
    s["maxiter"] = maxiter
    return torch.from_numpy(s.view(np.uint8).copy()).to(device)


def _read_state(t: torch.Tensor) -> np.void:
    return t.cpu().numpy().view(_STATE_DTYPE)[0]


class _PCGBuffers:
    """Device buffers of one CG solve (reused across solves of the same size)."""

    def __init__(self, n: int, maxiter: int, dtype, device):
        self.n = n
        self.maxiter = maxiter
        self.r = torch.empty(n, dtype=dtype, device=device)
        self.z = torch.empty(n, dtype=dtype, device=device)
        self.p = torch.empty(n, dtype=dtype, device=device)
        self.q = torch.empty(n, dtype=dtype, device=device)
        self.partials = torch.zeros(_lib.MAX_PARTIALS, dtype=torch.float64, device=device)
        self.log = torch.zeros((3, maxiter + 1), dtype=torch.float64, device=device)
        self.state = _state_tensor(0.0, maxiter, device)
        self.norm_out = torch.zeros(1, dtype=torch.float64, device=device)
        self.x = torch.zeros(n, dtype=dtype, device=device)
        self.b = torch.zeros(n, dtype=dtype, device=device)
        for v in (self.r, self.z, self.p, self.q):
            v.zero_()


class CG(SparseSolver):
    """Preconditioned conjugate gradients (solvers.py:192-303).

    Extra keyword arguments (no counterpart in the reference):
      check_every  iterations enqueued between two reads of the device-side state.
      use_graph    capture one block of `check_every` iterations in a hipGraph once and
                   replay it (single-RHS fast path only).
      spmv_format  "sell" (default) or "csr": the operator format for bsr_matmul=False.
    """

    def __init__(self, preconditioner: SparseSolver | None = None, rtol: float = 1e-7,
                 atol: float = 0, maxiter: int | None = 1000, bsr_matmul: bool = True,
                 estimate_cond: bool = False, check_every: int = 8, use_graph: bool = False,
                 spmv_format: str = "sell"):
        self.preconditioner = preconditioner
        self.rtol = rtol
        self.atol = atol
        self.maxiter = maxiter
        self.bsr_matmul = bsr_matmul
        self.estimate_cond = estimate_cond
        self.check_every = max(1, int(check_every))
        self.use_graph = use_graph
        if spmv_format not in ("sell", "csr"):
            raise ValueError(f"unknown spmv_format {spmv_format!r}")
        self.spmv_format = spmv_format
        self.matrix = None
        self._buf = None
        self._graph = None
        self._graph_tail = None

    def __str__(self):
        return f"CG({self.preconditioner or ''}, bsr_matmul={self.bsr_matmul})"

    def setup(self, matrix: torch.Tensor, *args, **kwargs) -> Metadata | None:
        meta = None
        if self.preconditioner is not None:
            meta = self.preconditioner.setup(matrix, *args, **kwargs)
        if self.bsr_matmul:
            if "block_size" in kwargs:
                block_size = kwargs["block_size"]
            else:
                raise ValueError("Unable to determine block size for BSR matrix multiplication.")
            self.matrix = FastBSR(csr_to_bsr(matrix, block_size=block_size))
        elif matrix.layout is torch.sparse_csr and matrix.values().is_cuda and \
                matrix.values().dtype in (torch.float32, torch.float64):
            self.matrix = _SELLOperator(matrix) if self.spmv_format == "sell" else _CSROperator(matrix)
        else:
            self.matrix = _TorchOperator(matrix)
        # captured graphs hold the addresses of the previous operator's arrays and buffers
        self._buf = None
        self._graph = None
        self._graph_tail = None
        self._bbuf = None
        self._bgraph = None
        return meta

    # -- helpers ----------------------------------------------------------------------
    def _fast(self) -> bool:
        ok_op = isinstance(self.matrix, (FastBSR, _CSROperator, _SELLOperator)) and \
            self.matrix.dtype == torch.float64
        ok_pc = self.preconditioner is None or getattr(self.preconditioner, "_device_apply", False)
        return ok_op and ok_pc

    def _norm(self, v: torch.Tensor, buf: _PCGBuffers) -> float:
        np_ = _lib.call_partials("dds_dot_partials", v.shape[0], v.data_ptr(), v.data_ptr(),
                                 buf.partials.data_ptr(), _lib.PARTIALS_OUT, None, _lib.stream())
        _lib.call("dds_sum_partials", buf.partials.data_ptr(), np_, buf.norm_out.data_ptr(),
                  _lib.stream())
        return float(buf.norm_out.item()) ** 0.5

    def _fin(self, op: int, count: int, buf: _PCGBuffers) -> None:
        lp = buf.log.data_ptr()
        stride = buf.log.stride(0) * 8
        _lib.call("dds_pcg_finalize", op, buf.partials.data_ptr(), count, buf.state.data_ptr(),
                  lp, lp + stride, lp + 2 * stride, _lib.stream())

    def _enqueue_iteration(self, x: torch.Tensor, buf: _PCGBuffers) -> None:
        st = buf.state.data_ptr()
        n = buf.n
        s = _lib.stream()
        if self.preconditioner is not None:
            cnt = self.preconditioner._enqueue_apply(buf.r, buf.z, buf.partials, buf.state)
            z = buf.z
        else:
            cnt = _lib.call_partials("dds_dot_partials", n, buf.r.data_ptr(), buf.r.data_ptr(),
                                     buf.partials.data_ptr(), _lib.PARTIALS_OUT, st, s)
            z = buf.r
        self._fin(_lib.FIN_RHO, cnt, buf)
        _lib.call("dds_pcg_update_p", n, z.data_ptr(), buf.p.data_ptr(), st, s)
        cnt = self.matrix._enqueue_matvec(buf.p, buf.q, buf.partials, buf.state)
        self._fin(_lib.FIN_ALPHA, cnt, buf)
        cnt = _lib.call_partials("dds_pcg_update_xr", n, buf.p.data_ptr(), buf.q.data_ptr(),
                                 x.data_ptr(), buf.r.data_ptr(), buf.partials.data_ptr(),
                                 _lib.PARTIALS_OUT, st, s)
        self._fin(_lib.FIN_NEXT, cnt, buf)

    def _metadata(self, norms, n_it, buf, alphas=None, betas=None) -> Metadata:
        meta = {
            "iterations": len(norms) - 1,
            "residual norms": norms,
            "preconditioner metadata": [{} for _ in range(len(norms) - 1)],
        }
        if self.estimate_cond:
            m = n_it
            T = np.zeros((m, m))
            for i in range(m):
                T[i, i] = 1.0 / alphas[i]
                if i > 0:
                    T[i, i] += betas[i] / alphas[i - 1]
                    T[i - 1, i] = T[i, i - 1] = betas[i] ** 0.5 / alphas[i - 1]
            if m:
                sv = np.linalg.svd(T, compute_uv=False)
                meta["condition number estimate"] = float(sv[0] / sv[-1])
            else:
                meta["condition number estimate"] = float("nan")
        return meta

    # -- solve ------------------------------------------------------------------------
    def solve(self, rhs: torch.Tensor, x0=None) -> Tuple[torch.Tensor, Metadata]:
        """Solve A x = rhs. A 2-D rhs (n, m) is a block of m independent systems solved
        together (each column with its own scalars and stopping test); the metadata then
        holds per-column lists."""
        if self.matrix is None:
            raise RuntimeError("CG.solve called before setup")
        _lib.require_device(rhs)
        with torch.cuda.device(rhs.device):
            pc = self.preconditioner
            # a single float64 vector takes the block path only where the local solver exists
            # for blocks alone (layout 'mfma')
            block_pc = pc is not None and getattr(pc, "_block_apply", False) and \
                getattr(pc, "_mfma", None) is not None
            if rhs.dim() == 2 or (block_pc and rhs.dim() == 1 and rhs.dtype == torch.float64):
                if not (self._fast_multi() and rhs.dtype == torch.float64):
                    raise ValueError("blocks of right-hand sides need float64 data, a SELL or CSR "
                                     "operator (bsr_matmul=False) and Inv(..., layout='mfma')")
                if rhs.dim() == 1:
                    x, meta = self._solve_block(rhs[:, None], None if x0 is None else x0[:, None])
                    meta1 = {k: (v[0] if isinstance(v, list) else v) for k, v in meta.items()}
                    if x0 is not None:
                        x0.copy_(x[:, 0])
                        return x0, meta1
                    return x[:, 0].clone(), meta1
                return self._solve_block(rhs, x0)
            if self._fast() and rhs.dtype == torch.float64 and rhs.dim() == 1:
                return self._solve_device(rhs.contiguous(), x0)
            return self._solve_generic(rhs, x0)

    # -- blocks of 16 right-hand sides ------------------------------------------------
    def _fast_multi(self) -> bool:
        pc = self.preconditioner
        return isinstance(self.matrix, _SELLOperator) and (pc is None or getattr(pc, "_block_apply", False))

    def _solve_block(self, rhs, x0):
        n, m = rhs.shape
        M = _lib.NRHS
        xs, metas = [], []
        for c0 in range(0, m, M):
            w = min(M, m - c0)
            xb, mb = self._solve_block16(rhs[:, c0:c0 + w], None if x0 is None else x0[:, c0:c0 + w], w)
            xs.append(xb)
            metas.append(mb)
        x = torch.cat(xs, dim=1) if len(xs) > 1 else xs[0]
        meta = {k: sum((mm[k] for mm in metas), []) for k in metas[0]}
        if x0 is not None:
            x0.copy_(x)
            x = x0
        if any(len(norms) == self._block_maxiter for norms in meta["residual norms"]):
            raise ConvergenceNotAchieved(meta)
        return x, meta

    def _solve_block16(self, rhs, x0, w):
        n = rhs.shape[0]
        M = _lib.NRHS
        dev = rhs.device
        maxiter = self.maxiter or n * 10
        self._block_maxiter = maxiter
        key = (n, maxiter)
        buf = getattr(self, "_bbuf", None)
        if buf is None or buf["key"] != key:
            z = lambda: torch.zeros((n, M), dtype=torch.float64, device=dev)  # noqa: E731
            buf = self._bbuf = dict(key=key, x=z(), r=z(), zz=z(), p=z(), q=z(), b=z(),
                                    partials=torch.zeros(_lib.MAX_PARTIALS * M, dtype=torch.float64, device=dev),
                                    log=torch.zeros((3, maxiter + 1, M), dtype=torch.float64, device=dev),
                                    state=torch.zeros(M * 72, dtype=torch.uint8, device=dev))
            self._bgraph = None
        B, X = buf["b"], buf["x"]
        B.zero_()
        B[:, :w] = rhs
        norms_b = torch.linalg.norm(B, dim=0).cpu().numpy()  # one host read, as solvers.py:229
        st = np.zeros(M, dtype=_STATE_DTYPE)
        st["atol"] = np.maximum(self.atol, self.rtol * norms_b)
        st["maxiter"] = maxiter
        st["done"][w:] = 3  # padding columns never run
        buf["state"].copy_(torch.from_numpy(st.view(np.uint8).copy()).to(dev))
        X.zero_()
        if x0 is not None:
            X[:, :w] = x0
        pc = self.preconditioner
        if pc is not None and hasattr(pc, "T0_multi"):
            X += pc.T0_multi(B)
        s = _lib.stream()
        self.matrix._enqueue_matmat(X, buf["q"], None, None)
        cnt = _lib.call_partials("dds_pcg_residual_m", n, B.data_ptr(), buf["q"].data_ptr(),
                                 buf["r"].data_ptr(), buf["partials"].data_ptr(), _lib.PARTIALS_OUT, s)
        self._fin_m(_lib.FIN_CHECK, cnt, buf)
        done_view = buf["state"].view(M, 72)[:, 60:64]
        enqueued = 0
        while enqueued < maxiter:
            if self.use_graph:
                if self._bgraph is None:
                    self._bgraph = self._capture_fn(lambda: [self._block_iteration(buf) for _ in range(self.check_every)])
                self._bgraph.replay()
                enqueued += self.check_every
            else:
                k = min(self.check_every, maxiter - enqueued)
                for _ in range(k):
                    self._block_iteration(buf)
                enqueued += k
            if bool((done_view.contiguous().view(torch.int32) != 0).all().item()):
                break
        stv = buf["state"].cpu().numpy().view(_STATE_DTYPE)
        log = buf["log"].cpu().numpy()
        meta = {"iterations": [], "residual norms": [], "preconditioner metadata": []}
        if self.estimate_cond:
            meta["condition number estimate"] = []
        for j in range(w):
            it, done = int(stv["it"][j]), int(stv["done"][j])
            nn = it + 1 if done == 1 else it
            mj = self._metadata([float(v) for v in log[0, :nn, j]], it, None,
                                log[1, :, j], log[2, :, j])
            for key in meta:
                meta[key].append(mj[key])
        return X[:, :w].clone(), meta

    def _fin_m(self, op, count, buf):
        lp = buf["log"].data_ptr()
        stride = buf["log"].stride(0) * 8
        _lib.call("dds_pcg_finalize_m", op, buf["partials"].data_ptr(), count,
                  buf["state"].data_ptr(), lp, lp + stride, lp + 2 * stride, _lib.stream())

    def _block_iteration(self, buf):
        n = buf["x"].shape[0]
        s = _lib.stream()
        st = buf["state"].data_ptr()
        pc = self.preconditioner
        if pc is not None:
            cnt = pc._enqueue_apply_multi(buf["r"], buf["zz"], buf["partials"], buf["state"])
            z = buf["zz"]
        else:
            cnt = _lib.call_partials("dds_dot_m", n, buf["r"].data_ptr(), buf["r"].data_ptr(),
                                     buf["partials"].data_ptr(), _lib.PARTIALS_OUT, st, s)
            z = buf["r"]
        self._fin_m(_lib.FIN_RHO, cnt, buf)
        _lib.call("dds_pcg_update_p_m", n, z.data_ptr(), buf["p"].data_ptr(), st, s)
        cnt = self.matrix._enqueue_matmat(buf["p"], buf["q"], buf["partials"], buf["state"])
        self._fin_m(_lib.FIN_ALPHA, cnt, buf)
        cnt = _lib.call_partials("dds_pcg_update_xr_m", n, buf["p"].data_ptr(), buf["q"].data_ptr(),
                                 buf["x"].data_ptr(), buf["r"].data_ptr(), buf["partials"].data_ptr(),
                                 _lib.PARTIALS_OUT, st, s)
        self._fin_m(_lib.FIN_NEXT, cnt, buf)

    @staticmethod
    def _capture_fn(fn):
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        return g

    def _solve_device(self, rhs, x0):
        n = rhs.shape[0]
        # BSR with b not dividing n pads the operator; the PCG vectors carry a zero tail
        # (empty padded rows keep it zero through every update).
        n_pad = max(n, int(self.matrix.shape[0]))
        maxiter = self.maxiter or n * 10
        buf = self._buf
        if buf is None or buf.n != n_pad or buf.maxiter != maxiter:
            buf = self._buf = _PCGBuffers(n_pad, maxiter, rhs.dtype, rhs.device)
            self._graph = None
        s = _lib.stream()
        b = rhs
        if n_pad > n:
            buf.b[:n] = rhs
            b = buf.b
        atol = max(self.atol, self.rtol * self._norm(b, buf))  # solvers.py:229
        buf.state.copy_(_state_tensor(atol, maxiter, rhs.device))
        # working x: buf.x (the address a captured graph holds), copied in and out
        x = buf.x
        if x0 is not None:
            x[:n] = x0
        else:
            x.zero_()
        if self.preconditioner is not None and hasattr(self.preconditioner, "T0"):
            x[:n] += self.preconditioner.T0(rhs)
        self.matrix._enqueue_matvec(x, buf.q, None, None)
        cnt = _lib.call_partials("dds_pcg_residual", n_pad, b.data_ptr(), buf.q.data_ptr(),
                                 buf.r.data_ptr(), buf.partials.data_ptr(), _lib.PARTIALS_OUT,
                                 None, s)
        self._fin(_lib.FIN_CHECK, cnt, buf)
        done_flag = buf.state[60:64].view(torch.int32)
        enqueued = 0
        if self.use_graph:
            if self._graph is None:
                self._capture(x, buf)
            # Blocks of check_every iterations; once the last block's convergence rate puts
            # the end within half a block, blocks of check_every / 4 (fewer gated launches
            # after convergence, at one more state read per short block).
            tail = self._graph_tail is not None
            short, prev = False, None
            while enqueued < maxiter:
                if short:
                    self._graph_tail.replay()
                    enqueued += self.check_every // 4
                else:
                    self._graph.replay()
                    enqueued += self.check_every
                if not tail:
                    if int(done_flag.item()):
                        break
                    continue
                stv = _read_state(buf.state)
                if int(stv["done"]):
                    break
                it, rr = int(stv["it"]), float(stv["rr"])
                if prev is not None and it > prev[0] and 0.0 < rr < prev[1]:
                    per_it = 0.5 * np.log(rr / prev[1]) / (it - prev[0])  # log norm ratio
                    short = np.log(atol / np.sqrt(rr)) / per_it <= self.check_every / 2
                prev = (it, rr)
        else:
            while enqueued < maxiter:
                k = min(self.check_every, maxiter - enqueued)
                for _ in range(k):
                    self._enqueue_iteration(x, buf)
                enqueued += k
                if int(done_flag.item()):
                    break
        stv = _read_state(buf.state)
        it, done = int(stv["it"]), int(stv["done"])
        n_norms = it + 1 if done == 1 else it
        log = buf.log[:, :max(n_norms, 1)].cpu().numpy()
        norms = [float(v) for v in log[0, :n_norms]]
        meta = self._metadata(norms, it, buf,
                              log[1] if self.estimate_cond else None,
                              log[2] if self.estimate_cond else None)
        if x0 is not None:  # the reference mutates and returns x0 (solvers.py:232, :266)
            x0.copy_(x[:n])
            out = x0
        else:
            out = x[:n].clone()
        if n_norms == maxiter:  # the loop index reached maxiter - 1 (solvers.py:294-295)
            raise ConvergenceNotAchieved(meta)
        return out, meta

    def _capture(self, x, buf):
        """Capture `check_every` iterations on a side stream into a hipGraph, and (for
        check_every >= 4) check_every / 4 iterations into a second one for the last blocks."""
        def capture(k):
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            # Capture with the solve already converged would record nothing different: every
            # kernel checks the gate at run time, not at capture time.
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    for _ in range(k):
                        self._enqueue_iteration(x, buf)
            torch.cuda.current_stream().wait_stream(side)
            return g
        self._graph = capture(self.check_every)
        self._graph_tail = capture(self.check_every // 4) if self.check_every >= 4 else None

    def _solve_generic(self, rhs, x0):
        """Reference control flow with a host check per iteration, for preconditioners and
        operators outside this package (solvers.py:228-297)."""
        atol = max(self.atol, self.rtol * float(torch.linalg.norm(rhs).item()))
        n = len(rhs)
        maxiter = self.maxiter or n * 10
        x = x0 if x0 is not None else torch.zeros_like(rhs)
        if self.preconditioner is not None and hasattr(self.preconditioner, "T0"):
            x += self.preconditioner.T0(rhs)
        r = rhs - self.matrix @ x
        rho_prev, p = None, None
        norms, pmeta, alphas, betas = [], [], [], []
        i = -1
        for i in range(maxiter):
            res = float(torch.linalg.norm(r).item())
            norms.append(res)
            if res < atol:
                break
            z, m = self.preconditioner.solve(r) if self.preconditioner is not None else (r, {})
            pmeta.append(m)
            rho = torch.dot(r, z)
            if i > 0:
                beta = rho / rho_prev
                p.mul_(beta).add_(z)
            else:
                p = z.clone()
            q = self.matrix @ p
            alpha = rho / torch.dot(p, q)
            x.add_(alpha * p)
            r.sub_(alpha * q)
            rho_prev = rho
            if self.estimate_cond:
                betas.append(float(beta) if i > 0 else 0.0)
                alphas.append(float(alpha))
        meta = self._metadata(norms, len(alphas), None, alphas, betas)
        meta["preconditioner metadata"] = pmeta
        if i + 1 == maxiter:
            raise ConvergenceNotAchieved(meta)
        return x, meta

    def destroy(self) -> None:
        if self.preconditioner is not None:
            self.preconditioner.destroy()
            self.preconditioner = None
        self.matrix = None
        self._buf = None
        self._graph = None
        self._graph_tail = None
        self._bbuf = None
        self._bgraph = None


# =======================================================================================
# Local solvers
# =======================================================================================


class DenseBatchSolver(ABC):
    @abstractmethod
    def setup(self, matrix_batch: torch.Tensor):
        """matrix_batch has shape (batch_size, n, n)."""

    @abstractmethod
    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        pass


class CSRBatch(NamedTuple):
    nrows: torch.Tensor
    nnz: torch.Tensor
    values: torch.Tensor
    col_indices: torch.Tensor
    row_offsets: torch.Tensor


class SparseBatchSolver(ABC):
    @abstractmethod
    def setup(self, matrix_batch: CSRBatch):
        pass

    @abstractmethod
    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        pass


def local_csr_batch(sub_ptr: torch.Tensor, s: torch.Tensor, a: torch.Tensor,
                    b: torch.Tensor, v: torch.Tensor, dtype: torch.dtype) -> CSRBatch:
    """The local matrices A_i as a CSRBatch, the layout of the reference's
    `_construct_local_solvers_matrices_sparse` (solvers.py:660-719, kernels :484-538).

    Entry e of the matrix lies in subdomain s[e] at local row a[e], local column b[e].
    Per subdomain i: nrows[i] = k_i (int32), nnz[i] (int64); the values (in `dtype`) and
    int32 local columns of its rows follow one another, subdomain by subdomain, with the
    columns of a row ascending; row_offsets holds k_i + 1 int32 offsets per subdomain,
    counted from the subdomain's first entry (sum(k_i + 1) in all; n + n_s for the
    reference's equal subdomains). Works on any device (torch only)."""
    dev = sub_ptr.device
    sizes = sub_ptr[1:] - sub_ptr[:-1]
    n_s = sizes.shape[0]
    nk = int(sub_ptr[-1])
    k_max = int(sizes.max()) if n_s else 0
    row = sub_ptr[s] + a  # position in the stacked rows of all subdomains
    order = torch.sort(row * max(k_max, 1) + b, stable=True).indices
    counts = torch.bincount(row, minlength=nk)
    G = torch.zeros(nk + 1, dtype=torch.int64, device=dev)
    G[1:] = torch.cumsum(counts, 0)
    starts = torch.cumsum(sizes + 1, 0) - (sizes + 1)
    q = torch.arange(int(sizes.sum()) + n_s, device=dev)
    sid = torch.repeat_interleave(torch.arange(n_s, device=dev), sizes + 1)
    j = q - starts[sid]
    row_offsets = (G[sub_ptr[sid] + j] - G[sub_ptr[sid]]).to(torch.int32)
    return CSRBatch(nrows=sizes.to(torch.int32), nnz=G[sub_ptr[1:]] - G[sub_ptr[:-1]],
                    values=v[order].to(dtype), col_indices=b[order].to(torch.int32),
                    row_offsets=row_offsets)


class Inv(DenseBatchSolver):
    """Explicit local inverses applied by a batched GEMV (solvers.py:343-359).

    The inverse is computed in the dtype of the matrices handed to `setup` (the
    preconditioner precision) and stored in `precision`. Extra keywords (no counterpart
    in the reference; they matter inside the Schwarz preconditioners):
      accumulate  accumulator dtype of the local apply (default: the vector dtype, as the
                  reference); torch.float32 gives FP32 accumulation for FP64 vectors.
      scaled      store A_i^-1 / max|A_i^-1| per subdomain and rescale the result (keeps
                  FP16 storage in range).
      layout      "auto" (default): "batched" for the reference's subdomains (equal,
                  contiguous, non-overlapping) of at most 64 unknowns with the
                  accumulator of the vector type, "full" for other subdomains of at most
                  64 unknowns, else "sym". "batched": the reference's (n, k, k)
                  column-major array, applied by the gemv_strided_batched kernels
                  (dds_local_apply_batched). "sym": the lower triangle in 64 x 64 tiles,
                  each off-diagonal tile read once and applied twice
                  (dds_local_apply_sym); "sym_mfma" (FP16/BF16 storage, chosen by "auto"
                  above 64 unknowns with an FP32 accumulator): the same triangle in MFMA
                  operand order, both products on the matrix cores with FP32 accumulation
                  and x split into 16-bit terms (dds_local_apply_sym_mfma); "full": the
                  whole matrix column-major with a
                  16-byte leading dimension (dds_local_apply); "mfma" (FP16/BF16): the
                  whole matrix in MFMA operand order, applied to blocks of 16 right-hand
                  sides as a batched GEMM on the matrix cores (dds_mfma_apply; a single
                  right-hand side runs as a block of one).
      autotune    with layout="mfma": time the apply's launch shapes once after setup and
                  keep the fastest (dd_solvers.tuning.autotune_mfma).
    """

    def __init__(self, precision: torch.dtype | None = None, accumulate: torch.dtype | None = None,
                 scaled: bool = False, layout: str = "auto", autotune: bool = False):
        if layout not in ("auto", "batched", "sym", "sym_mfma", "full", "mfma"):
            raise ValueError(f"unknown layout {layout!r}")
        if layout in ("mfma", "sym_mfma") and precision not in (torch.float16, torch.bfloat16):
            raise ValueError(f"layout={layout!r} stores the inverses in float16 or bfloat16")
        self.precision = precision
        self.accumulate = accumulate
        self.scaled = scaled
        self.layout = layout
        self.autotune = autotune

    def __str__(self):
        return f"Inv({self.precision or ''})"

    def setup(self, matrix_batch: torch.Tensor):
        """Inverts the (n, k, k) batch on the GPU (torch.linalg.inv in the reference,
        solvers.py:350-352): Gauss-Jordan with partial pivoting, dds_inverse_batched. A
        singular matrix raises torch.linalg.LinAlgError, as torch.linalg.inv does."""
        inv = inverse_batched(matrix_batch).to(self.precision or matrix_batch.dtype)
        # column-major storage, as gemv_strided_batched requires
        self.matrix_inv = inv.transpose(1, 2).contiguous().transpose(1, 2)

    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        return gemv_strided_batched(self.matrix_inv,
                                    rhs_batch.reshape(self.matrix_inv.shape[0], -1))


def inverse_batched(matrix_batch: torch.Tensor) -> torch.Tensor:
    """Inverse of every matrix of a (n, k, k) float64/float32 GPU batch (k <= 4096), by
    Gauss-Jordan with partial pivoting (dds_inverse_batched); a singular matrix raises
    torch.linalg.LinAlgError."""
    _lib.require_device(matrix_batch)
    if matrix_batch.dim() != 3 or matrix_batch.shape[1] != matrix_batch.shape[2]:
        raise ValueError("inverse_batched: expected a (n, k, k) batch")
    if matrix_batch.dtype not in (torch.float64, torch.float32):
        raise TypeError(f"inverse_batched: float64 or float32 matrices (got {matrix_batch.dtype})")
    n, k, _ = matrix_batch.shape
    a = matrix_batch.contiguous().clone()
    if n == 0 or k == 0:
        return a
    ks = torch.full((n,), k, dtype=torch.int32, device=a.device)
    info = torch.zeros(n, dtype=torch.int32, device=a.device)
    with torch.cuda.device(a.device):
        _lib.call("dds_inverse_batched", _lib.dtype_code(a.dtype), n, ks.data_ptr(), k, k * k,
                  a.data_ptr(), info.data_ptr(), _lib.stream())
    bad = torch.nonzero(info).flatten()
    if bad.numel():
        b = int(bad[0])
        raise torch.linalg.LinAlgError(
            f"inverse_batched: matrix {b} is singular (no pivot in column {int(info[b]) - 1})")
    return a


class LU(DenseBatchSolver):
    """Local solves by batched LU without pivoting (solvers.py:362-373), torch.linalg."""

    def __str__(self):
        return "LU"

    def setup(self, matrix_batch: torch.Tensor):
        self.matrix_lu = torch.linalg.lu_factor(matrix_batch, pivot=False)
        self.n_solvers = matrix_batch.shape[0]

    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        return torch.linalg.lu_solve(*self.matrix_lu, rhs_batch.reshape(self.n_solvers, -1, 1))


class Cholesky(DenseBatchSolver):
    """Local solves by batched Cholesky (solvers.py:376-387), torch.linalg."""

    def __str__(self):
        return "Cholesky"

    def setup(self, matrix_batch: torch.Tensor):
        self.matrix_cholesky = torch.linalg.cholesky(matrix_batch)

    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        return torch.cholesky_solve(rhs_batch.reshape(self.matrix_cholesky.shape[0], -1, 1),
                                    self.matrix_cholesky)


# =======================================================================================
# Schwarz preconditioners
# =======================================================================================


@dataclass
class _LocalStore:
    """Local inverses packed for dds_local_apply (device tensors)."""

    mats: torch.Tensor      # flat storage, each A_i^-1 column-major with leading dim ld_i
    mat_off: torch.Tensor   # (n_s,) int64 element offsets
    ld: torch.Tensor        # (n_s,) int32
    scale: torch.Tensor | None  # (n_s,) float32
    tiles: torch.Tensor     # (n_tiles, 2) int32: (subdomain, first row)
    n_tiles: int
    max_k: int
    mat_dtype: torch.dtype
    acc_dtype: torch.dtype
    bytes_matrices: int
    layout: str = "full"
    items: torch.Tensor | None = None   # sym: (n_items, 2) int32 (subdomain, pair group)
    ypos: torch.Tensor | None = None    # sym: (n_s,) int64 offsets of the partial vectors
    n_items: int = 0
    group: int = 1                      # sym: tile-row pairs per item
    direct: bool = False                # sym_mfma: one item per subdomain, written into Y
    bytes_algorithmic: int = 0          # lower triangle (sym) or full matrix, storage dtype


_SYM_INDEX_CACHE: dict = {}


def _sym_index(k: int, V: int) -> np.ndarray:
    """Source of every element of the symmetric tile layout of a k x k matrix with 16-byte
    chunks of V elements (csrc/local_sym.hip): row * k + col, -1 for padding (zero), and
    -2 - (row * k + col) for a diagonal entry, which is stored halved. Off-diagonal tiles
    are whole 64-wide chunk columns; a diagonal tile keeps, in chunk column j, only the
    rows r >= jV, with the entries above the diagonal zero."""
    key = (k, V)
    if key in _SYM_INDEX_CACHE:
        return _SYM_INDEX_CACHE[key]
    nch_full = 64 // V
    T = -(-k // 64)
    rL = k - 64 * (T - 1)
    parts = []
    for I in range(T):
        R = 64 if I < T - 1 else rL
        for J in range(I + 1):
            nch = nch_full if J < T - 1 else -(-rL // V)
            if J < I:
                j = np.arange(nch)[:, None, None]
                r = np.arange(R)[None, :, None]
                e = np.arange(V)[None, None, :]
                row = 64 * I + r + 0 * j + 0 * e
                col = 64 * J + j * V + e + 0 * r
                parts.append(np.where(col < k, row * k + col, -1).reshape(-1))
                continue
            for j in range(nch):
                r = np.arange(j * V, R)[:, None]
                c = j * V + np.arange(V)[None, :]
                row, col = 64 * I + r, 64 * J + c
                flat = row * k + col
                out = np.where((col < k) & (c <= r), flat, -1)
                parts.append(np.where(c == r, -2 - flat, out).reshape(-1))
    idx = np.concatenate(parts).astype(np.int64)
    _SYM_INDEX_CACHE[key] = idx
    return idx


_SYMM_INDEX_CACHE: dict = {}


def _symm_index(k: int) -> np.ndarray:
    """Source of every element of the matrix-core symmetric layout of a k x k matrix
    (csrc/local_sym.hip, sym_mfma_apply_kernel), with the encoding of _sym_index. Tiles
    (I, J), J <= I, in order; each keeps its non-zero 16 x 32 blocks (Rb, K2) in that order,
    lane l of a block holding T[16 Rb + l%16][32 K2 + 8(l/16) + e], e < 8. Diagonal tiles
    hold their lower triangle with the diagonal halved and skip the blocks above it."""
    if k in _SYMM_INDEX_CACHE:
        return _SYMM_INDEX_CACHE[k]
    T = -(-k // 64)
    rL = k - 64 * (T - 1)
    lane = np.arange(64)[:, None]
    e = np.arange(8)[None, :]
    parts = []
    for I in range(T):
        R = 64 if I < T - 1 else rL
        nb = -(-R // 16)
        for J in range(I + 1):
            diag = J == I
            nk = -(-R // 32) if diag else 2
            for Rb in range(nb):
                for K2 in range(nk):
                    if diag and K2 == 1 and Rb < 2:
                        continue
                    r = 16 * Rb + lane % 16 + 0 * e
                    c = 32 * K2 + 8 * (lane // 16) + e
                    row, col = 64 * I + r, 64 * J + c
                    flat = row * k + col
                    ok = (row < k) & (col < k) & (~diag | (c <= r))
                    out = np.where(ok, flat, -1)
                    if diag:
                        out = np.where(ok & (c == r), -2 - flat, out)
                    parts.append(out.reshape(-1))
    idx = np.concatenate(parts).astype(np.int64)
    _SYMM_INDEX_CACHE[k] = idx
    return idx


def _sym_pair_tasks(T: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Tiles that tile-row pair p (rows p and T-1-p) of a subdomain of T tile rows reads."""
    return np.where(p == T - 1 - p, T - p, T + 1)


def _sym_plan(sizes: np.ndarray, group: int = 1):
    """Work items (subdomain, pair group q) and the offsets of their partial vectors for the
    symmetric layouts: subdomain s has P_s = ceil(T_s / 2) tile-row pairs (pair p covers
    tile rows p and T_s-1-p); item q covers pairs q*group .. q*group + group - 1 and writes
    a 64*T_s-long partial vector (its first 64 (T_s - q*group) rows)."""
    T = -(-np.asarray(sizes, dtype=np.int64) // 64)
    P = -(-T // 2)
    Q = -(-P // group)
    ypos = np.concatenate([[0], np.cumsum(Q * 64 * T)]).astype(np.int64)
    items = _enumerate(Q)
    # longest items first (tiles read), so the last workgroups to start are the short ones
    Ts = T[items[:, 0]]
    work = np.zeros(items.shape[0], dtype=np.int64)
    for j in range(group):
        p = items[:, 1].astype(np.int64) * group + j
        work += np.where(p < -(-Ts // 2), _sym_pair_tasks(Ts, p), 0)
    return items[np.argsort(-work, kind="stable")], ypos


def _symm_group(sizes: np.ndarray, n_cu: int) -> int:
    """Tile-row pairs per item of the matrix-core symmetric apply. An item sums its products
    in LDS and writes each row it touches once, so larger items write less (whole subdomains:
    exactly the result, no combine) and start less often, while more items balance the CUs
    better: whole subdomains when every CU gets at least 8 of them, otherwise the largest
    power of two that keeps 32 items per CU (items of one pair at most)."""
    T = -(-np.asarray(sizes, dtype=np.int64) // 64)
    P = -(-T // 2)
    if sizes.shape[0] >= 8 * n_cu:
        return int(P.max())
    g = 1
    while 2 * g <= P.max() and int((-(-P // (2 * g))).sum()) >= 32 * n_cu:
        g *= 2
    return g


def _sym_written_rows(sizes: np.ndarray, group: int = 1) -> int:
    """Entries of the partial vectors that the symmetric apply writes (and its combine
    reads): item q writes the valid rows of blocks 0 .. T-1-q*group."""
    total = 0
    for k, cnt in zip(*np.unique(np.asarray(sizes, dtype=np.int64), return_counts=True)):
        k, T = int(k), -(-int(k) // 64)
        P = -(-T // 2)
        total += int(cnt) * sum(min(64 * (T - q * group), k) for q in range(-(-P // group)))
    return total


_MFMA_INDEX_CACHE: dict = {}
_SYM16_MAX_K = 1024  # largest subdomain of the 16-column symmetric apply (csrc/local_sym.hip)


def _mfma_index(k: int) -> np.ndarray:
    """Source index (row * k + col, -1 for padding) of every element of the MFMA A-operand
    layout of a k x k matrix (csrc/multi.hip): [row block R][column pair K2][lane l][8],
    lane l holding row 16R + l%16, columns 32K2 + 4(l/16) + i then 32K2 + 16 + 4(l/16) + i."""
    if k in _MFMA_INDEX_CACHE:
        return _MFMA_INDEX_CACHE[k]
    RB, KP = -(-k // 16), -(-k // 32)
    R = np.arange(RB)[:, None, None, None]
    K2 = np.arange(KP)[None, :, None, None]
    lane = np.arange(64)[None, None, :, None]
    e = np.arange(8)[None, None, None, :]
    row = 16 * R + lane % 16 + 0 * K2 + 0 * e
    col = 32 * K2 + 16 * (e // 4) + 4 * (lane // 16) + e % 4 + 0 * R
    idx = np.where((row < k) & (col < k), row * k + col, -1).reshape(-1).astype(np.int64)
    _MFMA_INDEX_CACHE[k] = idx
    return idx


def _enumerate(counts: np.ndarray) -> np.ndarray:
    """(s, j) for j < counts[s], s in order: an (sum(counts), 2) int32 array."""
    counts = np.asarray(counts, dtype=np.int64)
    s = np.repeat(np.arange(counts.shape[0], dtype=np.int64), counts)
    start = np.cumsum(counts) - counts
    j = np.arange(s.shape[0], dtype=np.int64) - np.repeat(start, counts)
    return np.stack([s, j], 1).astype(np.int32).reshape(-1, 2)


def _tiles_for(sizes: np.ndarray, rows: int) -> np.ndarray:
    t = _enumerate(-(-np.asarray(sizes, dtype=np.int64) // rows))
    t[:, 1] *= rows
    return t


class SchwarzOperator(SparseSolver):
    """Common setup of the Schwarz preconditioners (solvers.py:554-771).

    setup(matrix, dofs_per_solver, solvers_per_coarse, *, decomposition=None, waves=None)

    With `decomposition=None` the unknowns are already numbered so that solver i owns
    [i*k, (i+1)*k) and coarse cell c owns solvers_per_coarse[c] consecutive solvers (the
    reference's case). A `Decomposition` (dd_solvers.problems) or a mapping with
    owner_ptr / sub_ptr / sub_idx / solvers_per_coarse describes uneven, overlapping
    subdomains; `dofs_per_solver` is then ignored.
    """

    def __init__(self, preconditioner_precision: torch.dtype, local_solver, coarse_solver,
                 collect_timings: bool = False):
        self.preconditioner_precision = preconditioner_precision
        self.local_solver = local_solver
        self.coarse_solver = coarse_solver
        self.collect_timings = collect_timings
        self.waves = None
        self.fused_pack = True  # FP16 symmetric matrix-core apply: the x pack inside its kernel
        self._mfma = None
        self._msym = None

    # -- setup ------------------------------------------------------------------------
    def setup(self, matrix: torch.Tensor, dofs_per_solver: int | None = None,
              solvers_per_coarse: torch.Tensor | None = None, *args, decomposition=None,
              waves: int | None = None, **kwargs):
        if matrix.layout is not torch.sparse_csr:
            raise AssertionError("SchwarzOperator.setup expects a sparse CSR matrix")
        _lib.require_device(matrix.values())
        self.device = matrix.device
        self.dtype_in = matrix.dtype
        pi = self.preconditioner_precision or matrix.dtype
        if pi not in (torch.float64, torch.float32):
            raise ValueError(f"preconditioner precision must be float64 or float32, not {pi}")
        self.pi = pi
        self.waves = waves or self.waves
        n = matrix.shape[0]
        self.n = n
        # scratch vectors sized for the previous operator (n, n_coarse) are dropped; the
        # apply allocates them again at the new sizes on first use
        self._xc = self._yc = self._ybuf = None
        self._mbuf = None
        self._cast_buf = None
        self._mfma = None
        self._msym = None
        self.Yp = None
        self._xpack = None
        if self.collect_timings:
            events = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
            events[0].record()

        if decomposition is None:
            if dofs_per_solver is None or solvers_per_coarse is None:
                raise ValueError("setup needs dofs_per_solver and solvers_per_coarse "
                                 "(or decomposition=)")
            spc = torch.as_tensor(solvers_per_coarse).to("cpu", torch.int64).numpy()
            k = int(dofs_per_solver)
            n_s = int(spc.sum())
            if n_s * k != n:
                raise ValueError(f"{n_s} solvers x {k} dofs != {n} unknowns")
            owner_ptr = np.arange(0, n + 1, k, dtype=np.int64)
            sub_ptr = owner_ptr.copy()
            sub_idx = None
        else:
            get = (lambda name: getattr(decomposition, name)) if not isinstance(decomposition, dict) \
                else (lambda name: decomposition[name])
            owner_ptr = np.asarray(get("owner_ptr"), dtype=np.int64)
            sub_ptr = np.asarray(get("sub_ptr"), dtype=np.int64)
            sub_idx = np.asarray(get("sub_idx"), dtype=np.int64)
            try:
                spc = get("solvers_per_coarse")
            except (AttributeError, KeyError):
                spc = None
            n_s = owner_ptr.shape[0] - 1
            spc = np.ones(n_s, dtype=np.int64) if spc is None else np.asarray(spc, dtype=np.int64)
            if owner_ptr[-1] != n or sub_idx.shape[0] != sub_ptr[-1]:
                raise ValueError("decomposition does not match the matrix size")
            if sub_idx.shape[0] == n and np.array_equal(sub_ptr, owner_ptr) and \
                    np.array_equal(sub_idx, np.arange(n)):
                sub_idx = None  # non-overlapping and contiguous
        self.n_solvers = n_s
        self.n_coarse = int(spc.shape[0])
        self.solvers_per_coarse = torch.as_tensor(spc, device=self.device)
        scan = np.concatenate([[0], np.cumsum(spc)])
        self.solvers_per_coarse_scan = torch.as_tensor(scan.astype(np.int32), device=self.device)
        self.dofs_per_solver = int(dofs_per_solver) if decomposition is None else None
        coarse_ptr = owner_ptr[scan]
        # lanes per coarse cell of the restriction R_C x: a workgroup for large cells, a
        # group of 8-64 lanes for the reference's small DG cells
        avg = n / max(self.n_coarse, 1)
        self.restrict_lanes = 0 if avg > 256 else 64 if avg > 32 else 32 if avg > 16 else \
            16 if avg > 8 else 8
        sizes = np.diff(sub_ptr)
        self._host = dict(owner_ptr=owner_ptr, sub_ptr=sub_ptr, sub_idx=sub_idx,
                          coarse_ptr=coarse_ptr, sizes=sizes)
        dev = self.device
        self.sub_ptr = torch.as_tensor(sub_ptr, device=dev)
        self.sub_idx = None if sub_idx is None else torch.as_tensor(sub_idx.astype(np.int32), device=dev)
        self.coarse_ptr = torch.as_tensor(coarse_ptr, device=dev)
        cell = np.repeat(np.arange(self.n_coarse, dtype=np.int32), np.diff(coarse_ptr))
        self.cell = torch.as_tensor(cell, device=dev)
        if sub_idx is not None:
            order = np.argsort(sub_idx, kind="stable")
            counts = np.bincount(sub_idx, minlength=n)
            if counts.min() < 1:
                raise ValueError("every unknown must belong to at least one subdomain")
            self.dof_ptr = torch.as_tensor(np.concatenate([[0], np.cumsum(counts)]), device=dev)
            self.dof_pos = torch.as_tensor(order.astype(np.int32), device=dev)
            self.Y = torch.empty(int(sub_ptr[-1]), dtype=pi, device=dev)
        else:
            self.dof_ptr = self.dof_pos = self.Y = None

        if self.collect_timings:
            events[1].record()
        # Coarse matrix A_C = R_C A R_C^T (solvers.py:721-771), R0A = R_C A (:890-906).
        crow, col, val = matrix.crow_indices(), matrix.col_indices(), matrix.values()
        rows = csr_rows(crow)
        col64 = col.to(torch.int64)
        cell_t = self.cell.to(torch.int64)
        AC = self._segment_sum_csr(cell_t[rows] * self.n_coarse + cell_t[col64], val,
                                   self.n_coarse, self.n_coarse)
        collect()
        self.coarse_solver.setup(AC.to(pi))
        if self.collect_timings:
            events[2].record()
        self._setup_local(matrix, rows, col64, val, sizes, sub_idx)
        if self.collect_timings:
            events[3].record()
        self._setup_extra(matrix, rows, col64, val)
        self._cast_buf = None
        if self._mfma is not None and getattr(self.local_solver, "autotune", False):
            from .tuning import autotune_mfma
            self.tuning = autotune_mfma(self)
        if self.collect_timings:
            torch.cuda.synchronize()
            x_c = torch.rand(self.n_coarse, device=dev, dtype=pi)
            coarse_time = timeit(lambda: self.coarse_solver.solve(x_c))
            x = torch.rand(n, device=dev, dtype=pi)
            local_time = timeit(lambda: self._local_only(x))
            return {
                "preprocessing time": events[0].elapsed_time(events[1]),
                "coarse solver setup time": events[1].elapsed_time(events[2]),
                "local solver setup time": events[2].elapsed_time(events[3]),
                "coarse solver solve time": coarse_time,
                "local solver solve time": local_time,
            }
        return None

    def _setup_extra(self, matrix, rows, col64, val):
        pass

    @staticmethod
    def _segment_sum_csr(key: torch.Tensor, val: torch.Tensor, n_rows: int,
                         n_cols: int) -> torch.Tensor:
        """CSR (n_rows x n_cols, FP64) of the sums of val over equal keys row * n_cols + col,
        deterministic (stable sort, then compensated in-order segment sums: an entry of A_C
        sums a whole cell block of A, which nearly cancels); the coarse matrix stays sparse, so
        coarse spaces of millions of cells fit."""
        order = torch.sort(key, stable=True).indices
        ks = key[order]
        uk, counts = torch.unique_consecutive(ks, return_counts=True)
        sums = segment_sum(val[order].to(torch.float64), counts, compensated=True)
        crow = torch.zeros(n_rows + 1, dtype=torch.int64, device=val.device)
        crow[1:] = torch.cumsum(torch.bincount(uk // n_cols, minlength=n_rows), 0)
        return torch.sparse_csr_tensor(crow, uk % n_cols, sums, size=(n_rows, n_cols))

    def _setup_local(self, matrix, rows, col64, val, sizes, sub_idx):
        """Extract A_i = R_i A R_i^T, invert in the preconditioner precision and pack the
        inverses for dds_local_apply (solvers.py:645-658, :343-353)."""
        dev = self.device
        pi = self.pi
        n = self.n
        n_s = self.n_solvers
        k_max = int(sizes.max())
        ls = self.local_solver
        self._sparse_local = False
        if not isinstance(ls, (DenseBatchSolver, SparseBatchSolver)):
            raise ValueError(f"Unknown local solver type: {type(ls)}")
        sub_ptr = self.sub_ptr
        if sub_idx is None:
            # contiguous ranges: subdomain of an unknown by search in sub_ptr
            sub_r = torch.searchsorted(sub_ptr, rows, right=True) - 1
            sub_c = torch.searchsorted(sub_ptr, col64, right=True) - 1
            keep = sub_r == sub_c
            s = sub_r[keep]
            a = rows[keep] - sub_ptr[s]
            b = col64[keep] - sub_ptr[s]
            v = val[keep]
        else:
            # every membership (s, a) of the row, then the column looked up in S_s
            dof_ptr = self.dof_ptr
            mult = dof_ptr[1:] - dof_ptr[:-1]
            cnt = mult[rows]
            e = torch.repeat_interleave(torch.arange(rows.shape[0], device=dev), cnt)
            first = torch.cumsum(cnt, 0) - cnt
            t = torch.arange(e.shape[0], device=dev) - first[e]
            pos = self.dof_pos.to(torch.int64)[dof_ptr[rows[e]] + t]
            s = torch.searchsorted(sub_ptr, pos, right=True) - 1
            a = pos - sub_ptr[s]
            keys = torch.repeat_interleave(torch.arange(n_s, device=dev),
                                           sub_ptr[1:] - sub_ptr[:-1]) * n + self.sub_idx.to(torch.int64)
            qk = s * n + col64[e]
            q = torch.searchsorted(keys, qk).clamp_(max=keys.shape[0] - 1)
            found = keys[q] == qk
            s, a, b, v = s[found], a[found], (q - sub_ptr[s])[found], val[e][found]
            del e, first, t, pos, keys, qk, q, found, cnt, mult
        if isinstance(ls, SparseBatchSolver):
            # the reference's sparse local path (solvers.py:610-616): the A_i as a CSRBatch.
            # A solver with a device sweep (_enqueue_solve, BatchSparseCholesky) stays on the
            # gated device path, for any decomposition (it gathers overlapped subdomains
            # itself); other solvers see the reference's (n_s, k) batches of equal,
            # contiguous subdomains.
            device_path = hasattr(ls, "_enqueue_solve")
            if not device_path and (sub_idx is not None or not np.all(sizes == sizes[0])):
                raise ValueError(f"{ls} needs equal, non-overlapping subdomains")
            ls.setup(local_csr_batch(sub_ptr, s, a, b, v, pi))
            del s, a, b, v
            if device_path and self.sub_idx is not None:
                ls._gather_from(self.sub_idx)
            self.local = None
            self._generic_local = not device_path
            self._sparse_local = device_path
            collect()
            return
        if not isinstance(ls, Inv):
            # LU / Cholesky / user solvers: their own batched setup and solve on the
            # reference's layout (equal, contiguous, non-overlapping subdomains).
            if sub_idx is not None or not np.all(sizes == sizes[0]):
                raise ValueError(f"{ls} needs equal, non-overlapping subdomains; use Inv()")
            k = int(sizes[0])
            batch = torch.zeros(n_s * k * k, dtype=pi, device=dev)
            batch[(s * k + a) * k + b] = v.to(pi)
            ls.setup(batch.view(n_s, k, k))
            self.local = None
            self._generic_local = True
            collect()
            return
        self._generic_local = False
        store_dtype = ls.precision or pi
        acc = getattr(ls, "accumulate", None) or pi
        if store_dtype == torch.float64 and acc != torch.float64:
            raise ValueError("float64 local inverses need a float64 accumulator")
        width = _lib._lib.dds_local_vec_width(_lib.dtype_code(store_dtype))
        rows_per_tile = _lib._lib.dds_local_tile_rows(_lib.dtype_code(store_dtype))
        layout = getattr(ls, "layout", "full")
        reference_case = sub_idx is None and bool(np.all(sizes == sizes[0]))
        low16 = store_dtype in (torch.float16, torch.bfloat16)
        if layout == "auto":
            if reference_case and k_max <= 64 and acc == pi:
                layout = "batched"
            elif k_max <= 64:
                layout = "full"
            else:
                layout = "sym_mfma" if low16 and acc == torch.float32 and k_max <= 4096 else "sym"
        if layout == "sym_mfma" and not (low16 and acc == torch.float32):
            raise ValueError("layout='sym_mfma' needs float16/bfloat16 storage and a float32 accumulator")
        if layout == "batched" and not reference_case:
            raise ValueError("layout='batched' needs equal, contiguous, non-overlapping subdomains")
        if layout == "batched" and acc != pi:
            raise ValueError("layout='batched' accumulates in the vector dtype")
        if layout == "mfma":
            if store_dtype not in (torch.float16, torch.bfloat16):
                raise ValueError("layout='mfma' needs float16 or bfloat16 storage")
            mf_idx = {int(k): _mfma_index(int(k)) for k in np.unique(sizes)}
            ld = sizes.copy()
            mat_off = np.concatenate([[0], np.cumsum([mf_idx[int(k)].shape[0] for k in sizes])]).astype(np.int64)
            mf_idx_dev = {}
            acc = torch.float32
        elif layout in ("sym", "sym_mfma"):
            sym_idx = {int(k): (_sym_index(int(k), width) if layout == "sym" else _symm_index(int(k)))
                       for k in np.unique(sizes)}
            sym_len = np.array([sym_idx[int(k)].shape[0] for k in sizes], dtype=np.int64)
            ld = sizes.copy()
            mat_off = np.concatenate([[0], np.cumsum(sym_len)]).astype(np.int64)
            sym_idx_dev = {}
        else:
            # "batched": the reference's (n, k, k) column-major array (ld = k)
            ld = sizes.copy() if layout == "batched" else ((sizes + width - 1) // width) * width
            mat_off = np.concatenate([[0], np.cumsum(ld * sizes)]).astype(np.int64)
        total = int(mat_off[-1])
        mats = torch.zeros(total, dtype=store_dtype, device=dev)
        scaled = bool(getattr(ls, "scaled", False))
        scale = torch.ones(n_s, dtype=torch.float32, device=dev) if scaled else None
        # Invert in chunks of subdomains (bounded temporary memory).
        chunk = max(1, int(2 ** 31 // max(k_max * k_max * 8, 1)))
        flat_key = (s * k_max + a) * k_max + b
        order = torch.sort(flat_key).indices
        flat_key, v = flat_key[order], v[order]
        def build_chunk(c0, c1):
            lo = int(torch.searchsorted(flat_key, torch.tensor(c0 * k_max * k_max, device=dev)))
            hi = int(torch.searchsorted(flat_key, torch.tensor(c1 * k_max * k_max, device=dev)))
            blk = torch.zeros((c1 - c0) * k_max * k_max, dtype=pi, device=dev)
            blk[flat_key[lo:hi] - c0 * k_max * k_max] = v[lo:hi].to(pi)
            blk = blk.view(c1 - c0, k_max, k_max)
            kk = torch.as_tensor(sizes[c0:c1], device=dev)
            ar = torch.arange(k_max, device=dev)
            pad = ar[None, :] >= kk[:, None]
            blk.diagonal(dim1=1, dim2=2)[pad] = 1.0  # identity padding
            return blk, kk

        for c0 in range(0, n_s, chunk):
            c1 = min(n_s, c0 + chunk)
            blk, kk = build_chunk(c0, c1)
            # local matrices of an SPD operator are SPD: batched Gauss-Jordan inverse in
            # place; a chunk holding a matrix with a non-positive pivot is rebuilt and
            # inverted with partial pivoting (dds_inverse_batched; the reference's Inv.setup
            # uses torch.linalg.inv for every matrix)
            info = torch.zeros(c1 - c0, dtype=torch.int32, device=dev)
            _lib.call("dds_spd_inverse_batched", _lib.dtype_code(pi), c1 - c0,
                      kk.to(torch.int32).contiguous().data_ptr(), k_max, k_max * k_max,
                      blk.data_ptr(), info.data_ptr(), _lib.stream())
            if int(info.max()) == 0:
                inv = blk
            else:  # a local matrix that is not SPD: again from the start, with pivoting
                blk, kk = build_chunk(c0, c1)
                info.zero_()
                _lib.call("dds_inverse_batched", _lib.dtype_code(pi), c1 - c0,
                          kk.to(torch.int32).contiguous().data_ptr(), k_max, k_max * k_max,
                          blk.data_ptr(), info.data_ptr(), _lib.stream())
                bad = torch.nonzero(info).flatten()
                if bad.numel():
                    raise torch.linalg.LinAlgError(
                        f"local matrix of subdomain {c0 + int(bad[0])} is singular")
                inv = blk
            # pack every subdomain of one size at once
            ks = sizes[c0:c1]
            for k in np.unique(ks):
                k = int(k)
                js = np.flatnonzero(ks == k)
                jt = torch.as_tensor(js, device=dev)
                m = inv[jt, :k, :k]
                if scaled:
                    sc = m.abs().amax(dim=(1, 2)).to(torch.float32)
                    scale[jt + c0] = sc
                    m = m / sc.to(m.dtype)[:, None, None]
                if layout == "mfma":
                    if k not in mf_idx_dev:
                        mf_idx_dev[k] = torch.as_tensor(mf_idx[k], device=dev)
                    src = mf_idx_dev[k]
                    packed = torch.where(src >= 0, m.reshape(len(js), -1)[:, src.clamp(min=0)],
                                         torch.zeros((), dtype=pi, device=dev))
                elif layout in ("sym", "sym_mfma"):
                    if k not in sym_idx_dev:
                        sym_idx_dev[k] = torch.as_tensor(sym_idx[k], device=dev)
                    src = sym_idx_dev[k]
                    half = src <= -2
                    at = torch.where(half, -2 - src, src).clamp(min=0)
                    vals = m.reshape(len(js), -1)[:, at]
                    packed = torch.where(src >= 0, vals, torch.where(half, 0.5 * vals,
                                                                     torch.zeros((), dtype=pi, device=dev)))
                else:
                    # column-major with padded leading dimension: element (a, b) at b*ld + a
                    ldk = int(ld[c0 + js[0]])
                    col_major = torch.zeros((len(js), k, ldk), dtype=pi, device=dev)
                    col_major[:, :, :k] = m.transpose(1, 2)
                    packed = col_major.reshape(len(js), -1)
                off = mat_off[c0 + js]
                L = packed.shape[1]
                if np.all(np.diff(off) == L):  # consecutive subdomains: one slice
                    mats[int(off[0]):int(off[0]) + len(js) * L] = packed.reshape(-1).to(store_dtype)
                else:
                    dst = torch.as_tensor(off, device=dev)[:, None] + \
                        torch.arange(L, device=dev)[None, :]
                    mats[dst.reshape(-1)] = packed.reshape(-1).to(store_dtype)
                del m, packed
            del blk, inv
        esz = mats.element_size()
        if layout == "mfma":
            NT = 2 if store_dtype == torch.float16 else 3
            RB = -(-sizes // 16)
            KP = -(-sizes // 32)
            groups = -(-RB // 4)
            items = _enumerate(groups)
            # longest items first: work = row blocks in the group x column pairs
            work = np.minimum(4, RB[items[:, 0]] - 4 * items[:, 1]) * KP[items[:, 0]]
            items = items[np.argsort(-work, kind="stable")]
            b_off = np.concatenate([[0], np.cumsum(KP * NT * 512)]).astype(np.int64)
            tiles = np.zeros((0, 2), dtype=np.int32)
            ypos = None
            # one workgroup per subdomain (grid -1), longest first, is the default shape;
            # the autotuner also tries the persistent item grid (grid >= 0)
            order = np.argsort(-sizes, kind="stable").astype(np.int32)
            self._mfma = dict(
                b_off=torch.as_tensor(b_off[:-1], device=dev),
                bp=torch.empty(int(b_off[-1]), dtype=store_dtype, device=dev),
                sigma=torch.empty((n_s, _lib.NRHS), dtype=torch.float32, device=dev),
                Ym=torch.empty((int(sizes.sum()), _lib.NRHS), dtype=torch.float64, device=dev),
                order=torch.as_tensor(order, device=dev), grid=-1, waves=8, rpw=4)
            alg = int((sizes.astype(np.int64) ** 2).sum()) * esz
        elif layout in ("sym", "sym_mfma"):
            group = 1
            if layout == "sym_mfma":
                group = _symm_group(sizes, torch.cuda.get_device_properties(dev).multi_processor_count)
            items, ypos = _sym_plan(sizes, group)
            tiles = np.zeros((0, 2), dtype=np.int32)
            direct = layout == "sym_mfma" and items.shape[0] == n_s
            # one item per subdomain writes the result itself (ypos = sub_ptr)
            self.Yp = None if direct else torch.empty(int(ypos[-1]), dtype=pi, device=dev)
            if direct and self.Y is not None and pi == torch.float64:
                # the overlapped local results in FP32 (the precision they are accumulated
                # in): half the bytes written while the tiles stream in, summed in FP64
                self.Y = torch.empty(self.Y.shape[0], dtype=torch.float32, device=dev)
            if layout == "sym_mfma":  # per apply: x packed per 64-row block (scales, terms)
                nt = 2 if store_dtype == torch.float16 else 3
                T = -(-sizes.astype(np.int64) // 64)
                blk_ptr = np.concatenate([[0], np.cumsum(T)]).astype(np.int64)
                n_blk = int(blk_ptr[-1])
                self._xpack = dict(
                    n_blk=n_blk, blk_ptr=torch.as_tensor(blk_ptr, device=dev),
                    blk_owner=torch.as_tensor(np.repeat(np.arange(n_s, dtype=np.int32), T), device=dev),
                    sig=torch.ones(n_blk, dtype=torch.float32, device=dev),
                    xt=torch.zeros(n_blk * nt * 64, dtype=torch.int16, device=dev))
                if pi == torch.float64 and k_max <= _SYM16_MAX_K:
                    # blocks of 16 right-hand sides on the same tiles (dds_sym_mfma_apply16):
                    # X packed per (64-row block, column): scales and 16-bit terms
                    self._msym = dict(
                        bp=torch.zeros(n_blk * 2 * nt * 512, dtype=store_dtype, device=dev),
                        sigma=torch.empty((n_blk, _lib.NRHS), dtype=torch.float32, device=dev),
                        Ym=torch.empty((int(sizes.sum()), _lib.NRHS), dtype=torch.float64, device=dev),
                        order=torch.as_tensor(np.argsort(-sizes, kind="stable").astype(np.int32),
                                              device=dev))
            alg = int((sizes.astype(np.int64) * (sizes + 1) // 2).sum()) * esz
        else:
            tiles = _tiles_for(sizes, rows_per_tile)
            items = ypos = None
            alg = int((sizes.astype(np.int64) ** 2).sum()) * esz
        if layout != "mfma":
            self._mfma = None
        self.local = _LocalStore(
            mats=mats, mat_off=torch.as_tensor(mat_off[:-1], device=dev),
            ld=torch.as_tensor(ld.astype(np.int32), device=dev), scale=scale,
            tiles=torch.as_tensor(tiles, device=dev), n_tiles=int(tiles.shape[0]),
            max_k=k_max, mat_dtype=store_dtype, acc_dtype=acc,
            bytes_matrices=int(total * esz), layout=layout,
            items=None if items is None else torch.as_tensor(items, device=dev),
            ypos=None if ypos is None else torch.as_tensor(ypos[:-1], device=dev),
            n_items=0 if items is None else int(items.shape[0]), bytes_algorithmic=alg,
            group=group if layout in ("sym", "sym_mfma") else 1,
            direct=layout == "sym_mfma" and direct)
        collect()

    # -- apply pieces ---------------------------------------------------------------------
    @property
    def _device_apply(self) -> bool:
        """True when every step of the apply is a gated HIP kernel (CG's device loop)."""
        return (not getattr(self, "_generic_local", True)
                and hasattr(self.coarse_solver, "_enqueue_solve"))

    def _enqueue_local(self, x, Y, gate) -> None:
        if self._sparse_local:
            self.local_solver._enqueue_solve(x, Y, gate)
            return
        if self._generic_local:
            k = self.n // self.n_solvers
            Y[:self.n].copy_(self.local_solver.solve(x[:self.n].reshape(self.n_solvers, k)).reshape(-1))
            return
        L = self.local
        if L.layout == "mfma":
            raise RuntimeError("layout='mfma' applies blocks of right-hand sides (2-D); "
                               "CG routes single vectors through a block of one")
        if L.layout in ("sym", "sym_mfma"):
            self._enqueue_sym_apply(x, gate, Y)
            if not L.direct:
                self._enqueue_sym_combine(Y, gate)
            return
        if L.layout == "batched":
            k = L.max_k
            # LDS staging of x as gemv.py:8-14 chooses it
            lds = (5 < k) if self.pi == torch.float64 else (15 < k)
            _lib.call("dds_local_apply_batched", _lib.dtype_code(L.mat_dtype),
                      _lib.dtype_code(self.pi), L.mats.data_ptr(), x.data_ptr(), Y.data_ptr(),
                      self.n_solvers, k, int(lds), _lib.ptr(L.scale), _lib.ptr(gate),
                      _lib.stream())
            return
        _lib.call("dds_local_apply", _lib.dtype_code(L.mat_dtype), _lib.dtype_code(self.pi),
                  _lib.dtype_code(L.acc_dtype), L.n_tiles, L.tiles.data_ptr(),
                  self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx), L.max_k,
                  L.mat_off.data_ptr(), L.ld.data_ptr(), _lib.ptr(L.scale), L.mats.data_ptr(),
                  x.data_ptr(), Y.data_ptr(), self.waves or 4, _lib.ptr(gate), _lib.stream())

    def _enqueue_sym_pack(self, x, gate) -> None:
        """The x pack of the matrix-core symmetric apply (block scales and 16-bit terms)."""
        L, xp = self.local, self._xpack
        _lib.call("dds_sym_mfma_pack", _lib.dtype_code(L.mat_dtype), _lib.dtype_code(self.pi),
                  self.n_solvers, xp["n_blk"], xp["blk_owner"].data_ptr(), xp["blk_ptr"].data_ptr(),
                  self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx), x.data_ptr(),
                  xp["sig"].data_ptr(), xp["xt"].data_ptr(), _lib.ptr(gate), _lib.stream())

    def _enqueue_sym_apply_packed(self, gate, Y=None) -> None:
        """The matrix-core symmetric apply from the last pack of x."""
        L, xp = self.local, self._xpack
        out, pos = (Y, self.sub_ptr) if L.direct else (self.Yp, L.ypos)
        _lib.call("dds_sym_mfma_apply_packed", _lib.dtype_code(L.mat_dtype),
                  _lib.dtype_code(out.dtype), self.n_solvers, L.n_items, L.max_k, L.group,
                  L.items.data_ptr(), self.sub_ptr.data_ptr(), xp["blk_ptr"].data_ptr(),
                  L.mat_off.data_ptr(), pos.data_ptr(), _lib.ptr(L.scale), L.mats.data_ptr(),
                  xp["sig"].data_ptr(), xp["xt"].data_ptr(), out.data_ptr(), _lib.ptr(gate),
                  _lib.stream())

    def _enqueue_sym_apply(self, x, gate, Y=None) -> None:
        """Partial products of every (subdomain, pair group) item into Yp; for a direct
        layout (one item per subdomain) the result itself into Y."""
        L = self.local
        if L.layout == "sym_mfma":
            if L.direct and Y is None:
                raise ValueError("a direct sym_mfma apply writes the result: pass Y")
            out, pos = (Y, self.sub_ptr) if L.direct else (self.Yp, L.ypos)
            if L.mat_dtype == torch.float16 and self.fused_pack:  # one kernel, pack in LDS
                _lib.call("dds_local_apply_sym_mfma_fused", _lib.dtype_code(L.mat_dtype),
                          _lib.dtype_code(self.pi), _lib.dtype_code(out.dtype), self.n_solvers,
                          L.n_items, L.max_k, L.group, L.items.data_ptr(),
                          self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx),
                          L.mat_off.data_ptr(), pos.data_ptr(), _lib.ptr(L.scale),
                          L.mats.data_ptr(), x.data_ptr(), out.data_ptr(), _lib.ptr(gate),
                          _lib.stream())
                return
            if out.dtype != self.pi:  # FP32 results under FP64 vectors: the two halves
                self._enqueue_sym_pack(x, gate)
                self._enqueue_sym_apply_packed(gate, Y)
                return
            xp = self._xpack
            _lib.call("dds_local_apply_sym_mfma", _lib.dtype_code(L.mat_dtype),
                      _lib.dtype_code(self.pi), self.n_solvers, L.n_items, L.max_k, L.group,
                      L.items.data_ptr(), self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx),
                      L.mat_off.data_ptr(), pos.data_ptr(), _lib.ptr(L.scale),
                      L.mats.data_ptr(), x.data_ptr(), xp["n_blk"], xp["blk_owner"].data_ptr(),
                      xp["blk_ptr"].data_ptr(), xp["sig"].data_ptr(), xp["xt"].data_ptr(),
                      out.data_ptr(), _lib.ptr(gate), _lib.stream())
            return
        _lib.call("dds_local_apply_sym", _lib.dtype_code(L.mat_dtype), _lib.dtype_code(self.pi),
                  _lib.dtype_code(L.acc_dtype), L.n_items, L.items.data_ptr(),
                  self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx), L.mat_off.data_ptr(),
                  L.ypos.data_ptr(), _lib.ptr(L.scale), L.mats.data_ptr(), x.data_ptr(),
                  self.Yp.data_ptr(), _lib.ptr(gate), _lib.stream())

    def _enqueue_sym_combine(self, Y, gate) -> None:
        """Y[sub_ptr[s] + a] = sum of the partial vectors of subdomain s at row a."""
        L = self.local
        _lib.call("dds_sym_combine", _lib.dtype_code(self.pi), self.n_solvers, L.max_k, L.group,
                  self.sub_ptr.data_ptr(), L.ypos.data_ptr(), self.Yp.data_ptr(),
                  Y.data_ptr(), _lib.ptr(gate), _lib.stream())

    def _enqueue_overlap_sum(self, Y, z, yc, sign, r, partials, gate) -> int:
        overl = self.dof_ptr is not None and Y is self.Y
        return _lib.call_partials(
            "dds_overlap_sum", _lib.dtype_code(self.pi), self.n,
            _lib.ptr(self.dof_ptr) if overl else None, _lib.ptr(self.dof_pos) if overl else None,
            Y.data_ptr(), _lib.dtype_code(Y.dtype),
            _lib.ptr(self.cell) if yc is not None else None, _lib.ptr(yc),
            float(sign), z.data_ptr(), _lib.ptr(r), _lib.ptr(partials), _lib.PARTIALS_OUT,
            _lib.ptr(gate), _lib.stream())

    def _local_only(self, x):
        if self._mfma is not None:
            X = torch.zeros((self.n, _lib.NRHS), dtype=torch.float64, device=x.device)
            X[:, 0] = x
            return self._local_multi_into(X, None)[:, 0].to(x.dtype)
        y = torch.empty_like(x)
        if self.Y is None:
            self._enqueue_local(x, y, None)
        else:
            self._enqueue_local(x, self.Y, None)
            self._enqueue_overlap_sum(self.Y, y, None, 0.0, None, None, None)
        return y

    def _coarse_enqueue(self, xc, yc, gate) -> None:
        cs = self.coarse_solver
        if hasattr(cs, "_enqueue_solve"):
            cs._enqueue_solve(xc, yc, gate)
        else:
            out, _ = cs.solve(xc)
            yc.copy_(out)

    def _buffers(self):
        if getattr(self, "_xc", None) is None or self._xc.dtype != self.pi:
            self._xc = torch.empty(self.n_coarse, dtype=self.pi, device=self.device)
            self._yc = torch.empty(self.n_coarse, dtype=self.pi, device=self.device)
            self._ybuf = torch.empty(self.n, dtype=self.pi, device=self.device)
        return self._xc, self._yc, self._ybuf

    def _enqueue_apply(self, r, z, partials, gate) -> int:
        """z = M r on the current stream (every kernel gated); per-block partials of r . z
        are written to `partials`; returns their count. r, z are float64."""
        if self.pi == torch.float64:
            return self._enqueue_pi(r, z, r, partials, gate)
        if self._cast_buf is None:
            self._cast_buf = (torch.empty(self.n, dtype=self.pi, device=self.device),
                              torch.empty(self.n, dtype=self.pi, device=self.device))
        r32, z32 = self._cast_buf
        s = _lib.stream()
        _lib.call("dds_cast", _lib.F64, _lib.dtype_code(self.pi), self.n, r.data_ptr(),
                  r32.data_ptr(), _lib.ptr(gate), s)
        self._enqueue_pi(r32, z32, None, None, gate)
        _lib.call("dds_cast", _lib.dtype_code(self.pi), _lib.F64, self.n, z32.data_ptr(),
                  z.data_ptr(), _lib.ptr(gate), s)
        return _lib.call_partials("dds_dot_partials", self.n, r.data_ptr(), z.data_ptr(),
                                  partials.data_ptr(), _lib.PARTIALS_OUT, _lib.ptr(gate), s)

    # -- blocks of 16 right-hand sides ---------------------------------------------------
    @property
    def _block_apply(self) -> bool:
        """Blocks of 16 right-hand sides run on the matrix cores (layout 'mfma', or
        'sym_mfma' with k <= 1024)."""
        return (self._mfma is not None or self._msym is not None) and \
            self.pi == torch.float64 and hasattr(self.coarse_solver, "_enqueue_solve_multi")

    @property
    def _block_state(self) -> dict:
        return self._mfma if self._mfma is not None else self._msym

    def _enqueue_local_multi(self, X, Y, st) -> None:
        """Y[sub_ptr[s] + a][:] = (A_s^-1 X_s)[a][:] for (n, 16) X (MFMA)."""
        L = self.local
        s = _lib.stream()
        dt = _lib.dtype_code(L.mat_dtype)
        if self._mfma is None:  # the symmetric tiles
            ms, xp = self._msym, self._xpack
            _lib.call("dds_sym_mfma_pack16", dt, xp["n_blk"], xp["blk_owner"].data_ptr(),
                      xp["blk_ptr"].data_ptr(), self.sub_ptr.data_ptr(), _lib.ptr(self.sub_idx),
                      X.data_ptr(), ms["sigma"].data_ptr(), ms["bp"].data_ptr(), _lib.ptr(st), s)
            _lib.call("dds_sym_mfma_apply16", dt, self.n_solvers, L.max_k, ms["order"].data_ptr(),
                      self.sub_ptr.data_ptr(), xp["blk_ptr"].data_ptr(), L.mat_off.data_ptr(),
                      _lib.ptr(L.scale), ms["sigma"].data_ptr(), L.mats.data_ptr(),
                      ms["bp"].data_ptr(), Y.data_ptr(), _lib.ptr(st), s)
            return
        mf = self._mfma
        _lib.call("dds_mfma_pack_x", dt, self.n_solvers, self.sub_ptr.data_ptr(),
                  _lib.ptr(self.sub_idx), mf["b_off"].data_ptr(), X.data_ptr(),
                  mf["bp"].data_ptr(), mf["sigma"].data_ptr(), _lib.ptr(st), s)
        self._enqueue_mfma_apply(Y, st)

    def _enqueue_mfma_apply(self, Y, st) -> None:
        """The MFMA product of the packed right-hand sides into Y, in the launch shape of
        self._mfma: grid -1 one workgroup of `waves` wavefronts per subdomain taking `rpw`
        row blocks at a time (dds_mfma_apply_sub); grid >= 0 the persistent item grid
        (dds_mfma_apply, 0: one wavefront per item). Every shape computes the same bits."""
        L, mf = self.local, self._mfma
        dt = _lib.dtype_code(L.mat_dtype)
        s = _lib.stream()
        if int(mf["grid"]) < 0:
            _lib.call("dds_mfma_apply_sub", dt, self.n_solvers, mf["order"].data_ptr(),
                      self.sub_ptr.data_ptr(), L.mat_off.data_ptr(), mf["b_off"].data_ptr(),
                      _lib.ptr(L.scale), mf["sigma"].data_ptr(), L.mats.data_ptr(),
                      mf["bp"].data_ptr(), Y.data_ptr(), int(mf["waves"]), int(mf["rpw"]),
                      _lib.ptr(st), s)
        else:
            _lib.call("dds_mfma_apply", dt, L.n_items, L.items.data_ptr(), self.sub_ptr.data_ptr(),
                      L.mat_off.data_ptr(), mf["b_off"].data_ptr(), _lib.ptr(L.scale),
                      mf["sigma"].data_ptr(), L.mats.data_ptr(), mf["bp"].data_ptr(), Y.data_ptr(),
                      int(mf["grid"]), int(mf["waves"]), _lib.ptr(st), s)

    def _buffers_multi(self):
        if getattr(self, "_mbuf", None) is None:
            M = _lib.NRHS
            self._mbuf = dict(xc=torch.empty((self.n_coarse, M), dtype=torch.float64, device=self.device),
                              yc=torch.empty((self.n_coarse, M), dtype=torch.float64, device=self.device),
                              y=torch.empty((self.n, M), dtype=torch.float64, device=self.device))
        return self._mbuf

    def _overlap_sum_multi(self, Y, z, yc, sign, r, partials, st) -> int:
        overl = self.dof_ptr is not None and Y is self._block_state["Ym"]
        return _lib.call_partials(
            "dds_overlap_sum_m", self.n, _lib.ptr(self.dof_ptr) if overl else None,
            _lib.ptr(self.dof_pos) if overl else None, Y.data_ptr(),
            _lib.ptr(self.cell) if yc is not None else None, _lib.ptr(yc), float(sign),
            z.data_ptr(), _lib.ptr(r), _lib.ptr(partials), _lib.PARTIALS_OUT, _lib.ptr(st),
            _lib.stream())

    def _local_multi_into(self, X, st):
        """Local part for a block: returns the tensor holding it in unknown order."""
        Ym = self._block_state["Ym"]
        self._enqueue_local_multi(X, Ym, st)
        if self.dof_ptr is None:  # non-overlapping: Ym is already in unknown order
            return Ym
        y = self._buffers_multi()["y"]
        self._overlap_sum_multi(Ym, y, None, 0.0, None, None, st)
        return y

    def _enqueue_apply_multi(self, R, Z, partials, st) -> int:
        """Z = M R for an (n_pad, 16) block, gated per column; per-column partials of R . Z."""
        return self._enqueue_multi(R, Z, R, partials, st)

    def solve(self, rhs: torch.Tensor) -> Tuple[torch.Tensor, Metadata]:
        _lib.require_device(rhs)
        with torch.cuda.device(rhs.device):
            if rhs.dim() == 2:
                if not self._block_apply:
                    raise ValueError("2-D right-hand sides need Inv(..., layout='mfma')")
                out = torch.empty_like(rhs, dtype=torch.float64)
                M = _lib.NRHS
                for c0 in range(0, rhs.shape[1], M):
                    blk = torch.zeros((self.n, M), dtype=torch.float64, device=rhs.device)
                    w = min(M, rhs.shape[1] - c0)
                    blk[:, :w] = rhs[:, c0:c0 + w]
                    z = torch.empty_like(blk)
                    self._enqueue_multi(blk, z, None, None, None)
                    out[:, c0:c0 + w] = z[:, :w]
                return out.to(rhs.dtype), {}
            x = rhs.to(self.pi).contiguous()
            y = torch.empty_like(x)
            self._enqueue_pi(x, y, None, None, None)
        return y.to(rhs.dtype), {}

    def stored_local_inverses(self) -> list:
        """The local inverses as the device holds them: for each subdomain the k_i x k_i
        FP64 matrix of the stored values (storage rounding included) times the subdomain's
        scale, read back from the packed store. For checks against an exact product."""
        L = self.local
        if L is None:
            raise ValueError("no explicit local inverses (local solver is not Inv)")
        sizes = self._host["sizes"].astype(np.int64)
        mats = L.mats.to(torch.float64).cpu().numpy()
        off = L.mat_off.cpu().numpy()
        ld = L.ld.cpu().numpy()
        scale = L.scale.cpu().numpy().astype(np.float64) if L.scale is not None else None
        width = _lib._lib.dds_local_vec_width(_lib.dtype_code(L.mat_dtype))
        out = []
        for i, k in enumerate(sizes):
            k = int(k)
            o = int(off[i])
            if L.layout in ("sym", "sym_mfma"):
                idx = _sym_index(k, width) if L.layout == "sym" else _symm_index(k)
                vals = mats[o:o + idx.shape[0]]
                m = np.zeros(k * k)
                keep = idx >= 0
                m[idx[keep]] = vals[keep]
                half = idx <= -2
                m[-2 - idx[half]] = 2.0 * vals[half]
                m = m.reshape(k, k)
                m = np.tril(m) + np.tril(m, -1).T
            elif L.layout == "mfma":
                idx = _mfma_index(k)
                m = np.zeros(k * k)
                keep = idx >= 0
                m[idx[keep]] = mats[o:o + idx.shape[0]][keep]
                m = m.reshape(k, k)
            else:  # "full" / "batched": column-major, leading dimension ld
                l_ = int(ld[i])
                m = mats[o:o + l_ * k].reshape(k, l_)[:, :k].T.copy()
            out.append(m * (scale[i] if scale is not None else 1.0))
        return out

    def destroy(self) -> None:
        self.coarse_solver.destroy()
        self.local = None
        # a sparse local solver holds its factor on the device (BatchSparseCholesky.destroy);
        # the reference's local solvers have no destroy (solvers.py:642-643)
        if getattr(self, "_sparse_local", False):
            self.local_solver.destroy()

    @property
    def bytes_per_apply(self) -> int:
        """Algorithmic bytes of one application (matrices, vectors, coarse)."""
        e = 8 if self.pi == torch.float64 else 4
        sizes = self._host["sizes"]
        nk = int(sizes.sum())
        if self.local is not None:
            b = self.local.bytes_algorithmic + e * 2 * nk
        else:  # a sparse local factor: its sweeps' bytes, vectors included
            b = getattr(self.local_solver, "bytes_per_solve", 0)
        if self.sub_idx is not None:
            b += 4 * nk  # gather indices
        if getattr(self, "Yp", None) is not None:  # partial vectors: written, read back
            b += 2 * e * int(self.Yp.shape[0])
        if self.dof_pos is not None:  # overlapped sum: staged Y written and read with lists
            b += 2 * e * nk + 4 * nk + 8 * (self.n + 1)
        b += e * 2 * self.n  # r read by the restriction/sum, z written
        b += getattr(self.coarse_solver, "bytes_per_solve", 0)
        return b


class AdditiveSchwarz(SchwarzOperator):
    """M r = sum_i R_i^T A_i^-1 R_i r + R_C^T A_C^-1 R_C r (solvers.py:774-826)."""

    def __init__(self, preconditioner_precision: torch.dtype, local_solver, coarse_solver,
                 number_of_dofs_per_coarse_is_const: bool = False, collect_timings: bool = False):
        super().__init__(preconditioner_precision, local_solver, coarse_solver, collect_timings)
        self.number_of_dofs_per_coarse_is_const = number_of_dofs_per_coarse_is_const

    def __str__(self):
        return (f"AdditiveSchwarz({self.preconditioner_precision}, {self.local_solver}, "
                f"{self.coarse_solver}, const_dofs_per_coarse={self.number_of_dofs_per_coarse_is_const})")

    def setup(self, matrix, dofs_per_solver=None, solvers_per_coarse=None, *args, **kwargs):
        timings = super().setup(matrix, dofs_per_solver, solvers_per_coarse, *args, **kwargs)
        if self.number_of_dofs_per_coarse_is_const and self.dofs_per_solver is not None:
            self.dofs_per_coarse = self.dofs_per_solver * int(self.solvers_per_coarse[0].item())
        return timings

    def _enqueue_pi(self, x, z, r, partials, gate) -> int:
        xc, yc, _ = self._buffers()
        s = _lib.stream()
        Y = z if self.Y is None else self.Y
        self._enqueue_local(x, Y, gate)
        _lib.call("dds_restrict", _lib.dtype_code(self.pi), self.n_coarse,
                  self.coarse_ptr.data_ptr(), x.data_ptr(), xc.data_ptr(), self.restrict_lanes,
                  _lib.ptr(gate), s)
        self._coarse_enqueue(xc, yc, gate)
        return self._enqueue_overlap_sum(Y, z, yc, 1.0, r, partials, gate)


def _additive_multi(self, X, Z, R, partials, st) -> int:
    mb = self._buffers_multi()
    s = _lib.stream()
    Ym = self._block_state["Ym"]
    self._enqueue_local_multi(X, Ym, st)
    _lib.call("dds_restrict_m", self.n_coarse, self.coarse_ptr.data_ptr(), X.data_ptr(),
              mb["xc"].data_ptr(), _lib.ptr(st), s)
    self.coarse_solver._enqueue_solve_multi(mb["xc"], mb["yc"], st)
    return self._overlap_sum_multi(Ym, Z, mb["yc"], 1.0, R, partials, st)


AdditiveSchwarz._enqueue_multi = _additive_multi


class HybridSchwarz(SchwarzOperator):
    """M r = y - R_C^T A_C^-1 (R0A y), y = sum_i R_i^T A_i^-1 R_i r; T0 b = R_C^T A_C^-1 R_C b
    (solvers.py:829-906)."""

    def __init__(self, preconditioner_precision: torch.dtype, local_solver, coarse_solver,
                 collect_timings: bool = False):
        super().__init__(preconditioner_precision, local_solver, coarse_solver, collect_timings)

    def __str__(self):
        return f"HybridSchwarz({self.preconditioner_precision}, {self.local_solver}, {self.coarse_solver})"

    def _setup_extra(self, matrix, rows, col64, val):
        # R0A = R_C A: rows of each coarse cell merged, duplicate columns summed (:890-906)
        cell_t = self.cell.to(torch.int64)
        key = cell_t[rows] * self.n + col64
        order = torch.sort(key, stable=True).indices
        ks = key[order]
        uk, counts = torch.unique_consecutive(ks, return_counts=True)
        sums = segment_sum(val[order].to(torch.float64), counts, compensated=True)
        r_of = uk // self.n
        crow = torch.zeros(self.n_coarse + 1, dtype=torch.int64, device=self.device)
        crow[1:] = torch.cumsum(torch.bincount(r_of, minlength=self.n_coarse), 0)
        R0A = torch.sparse_csr_tensor(crow, (uk % self.n).to(torch.int32), sums.to(self.pi),
                                      size=(self.n_coarse, self.n))
        self.R0A = R0A
        self._r0a = _CSROperator(R0A)

    def T0(self, rhs: torch.Tensor) -> torch.Tensor:
        with torch.cuda.device(rhs.device):
            x = rhs.to(self.pi).contiguous()
            xc, yc, _ = self._buffers()
            s = _lib.stream()
            _lib.call("dds_restrict", _lib.dtype_code(self.pi), self.n_coarse,
                      self.coarse_ptr.data_ptr(), x.data_ptr(), xc.data_ptr(), self.restrict_lanes,
                      None, s)
            self._coarse_enqueue(xc, yc, None)
            zero = torch.zeros_like(x)
            y = torch.empty_like(x)
            self._enqueue_overlap_sum(zero, y, yc, 1.0, None, None, None)
        return y.to(rhs.dtype)

    def _enqueue_multi(self, X, Z, R, partials, st) -> int:
        mb = self._buffers_multi()
        y = self._local_multi_into(X, st)
        _lib.call("dds_csrmm_rows", self.n_coarse, self._r0a.crow.data_ptr(),
                  self._r0a.col.data_ptr(), self._r0a.values.data_ptr(), y.data_ptr(),
                  mb["xc"].data_ptr(), _lib.ptr(st), _lib.stream())
        self.coarse_solver._enqueue_solve_multi(mb["xc"], mb["yc"], st)
        return self._overlap_sum_multi(y, Z, mb["yc"], -1.0, R, partials, st)

    def T0_multi(self, B: torch.Tensor) -> torch.Tensor:
        """T0 for an (n, 16) block: R_C^T A_C^-1 R_C B."""
        mb = self._buffers_multi()
        s = _lib.stream()
        _lib.call("dds_restrict_m", self.n_coarse, self.coarse_ptr.data_ptr(), B.data_ptr(),
                  mb["xc"].data_ptr(), None, s)
        self.coarse_solver._enqueue_solve_multi(mb["xc"], mb["yc"], None)
        zero = torch.zeros_like(B)
        out = torch.empty_like(B)
        _lib.call_partials("dds_overlap_sum_m", self.n, None, None, zero.data_ptr(),
                           self.cell.data_ptr(), mb["yc"].data_ptr(), 1.0, out.data_ptr(), None,
                           None, _lib.PARTIALS_OUT, None, s)
        return out

    def _enqueue_pi(self, x, z, r, partials, gate) -> int:
        xc, yc, ybuf = self._buffers()
        if self.Y is None:
            self._enqueue_local(x, ybuf, gate)
        else:
            self._enqueue_local(x, self.Y, gate)
            self._enqueue_overlap_sum(self.Y, ybuf, None, 0.0, None, None, gate)
        self._r0a._enqueue_matvec(ybuf, xc, None, gate)
        self._coarse_enqueue(xc, yc, gate)
        return self._enqueue_overlap_sum(ybuf, z, yc, -1.0, r, partials, gate)

    @property
    def bytes_per_apply(self) -> int:
        return super().bytes_per_apply + self._r0a.bytes_per_matvec
