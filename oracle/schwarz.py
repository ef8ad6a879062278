"""Oracle restatement of the Schwarz preconditioners (numpy/scipy, CPU).

Follows src/dd_solvers/dd_solvers/solvers.py:
  SchwarzOperator.setup               :571-640  (scan of solvers_per_coarse, A_C, A_i)
  _construct_local_solvers_matrices_dense :645-658 (A_i = diagonal blocks of A)
  _construct_coarse_solver_matrix     :721-771  (A_C = R_C A R_C^T)
  Inv.setup / Inv.solve               :343-359  (inverse in the preconditioner precision,
                                                 then cast to the storage precision)
  AdditiveSchwarz.solve               :811-826  (y = sum R_i^T A_i^-1 R_i x + R_C^T A_C^-1 R_C x)
  HybridSchwarz.T0 / solve            :861-888  (T0 b = R_C^T A_C^-1 R_C b;
                                                 y - R_C^T A_C^-1 (R0A y), y = sum R_i^T A_i^-1 R_i x)
  _construct_R0A_matrix               :890-906  (R0A = R_C A)

Generalisation (no counterpart in the reference, which assumes equal, contiguous,
non-overlapping subdomains, solvers.py:554-556): subdomain i is any sorted index set
S_i (core plus overlap), R_i restricts to S_i, and the overlapped sum sum_i R_i^T y_i is
taken in ascending subdomain order. R_C sums over the coarse cells' unknown ranges (unions
of subdomain cores). For equal, non-overlapping subdomains this is exactly the reference.

Precision: `precision` is the reference's preconditioner_precision (arithmetic and the
dtype of A_i, A_C, R0A); `storage` is Inv.precision; `accumulate` is the local-apply
accumulator (the reference accumulates in the vector dtype, gemv_kernels.cu:47-54);
`scaled` stores A_i^-1 / max|A_i^-1| per subdomain (no counterpart in the reference).

`BlockSchwarzOracle` is the reference's own case (equal, contiguous, non-overlapping
subdomains of k unknowns, coarse cells = runs of `solvers_per_coarse` subdomains,
solvers.py:554-556, 590-598) in FP64, vectorised over the subdomains and with a sparse LU
coarse solve, so that the reference's DG runs (10^5 - 10^6 coarse cells) fit on a CPU.

Pins (tests/test_oracle.py): the reference's 24-unknown DG fixture solved to
||A x - b|| < 1e-10 (src/dd_solvers/tests/test_solvers.py:153-176); M_add built as an
explicit dense matrix is SPD; the Hybrid projection P = I - R_C^T A_C^-1 R_C A satisfies
P^2 = P and A P = P^T A (preprint.pdf p.6-7); one subdomain covering every unknown gives
M = A^-1 (one CG iteration); single-unknown subdomains without a coarse space give Jacobi.
BlockSchwarzOracle agrees with SchwarzOracle (tests/test_oracle.py) and, through
tests/test_dg.py, reproduces the iteration counts and residual histories of the
reference's published FP64 DG runs (results/experiment_precisions.csv).
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from .kernels import quantize

__all__ = ["SchwarzOracle", "BlockSchwarzOracle"]

_NP = {"float64": np.float64, "float32": np.float32}


class SchwarzOracle:
    def __init__(self, A: sps.csr_matrix, owner_ptr, sub_ptr, sub_idx, solvers_per_coarse=None,
                 kind: str = "additive", precision: str = "float64", storage: str | None = None,
                 accumulate: str | None = None, scaled: bool = False, coarse: bool = True,
                 local_subset=None, inverses=None):
        """local_subset: indices of the subdomains whose inverses are formed (default all).
        With a subset, local() sums only those subdomains: a timing sample of the local
        apply on a workload too large to invert whole on a host, never a preconditioner.
        inverses: the k_i x k_i matrices to apply as the local inverses, as given (no
        inversion, storage rounding or scaling here): the exact product of a solver's own
        stored inverses, against which its kernels' arithmetic alone is measured."""
        self.kind = kind
        self.pi = _NP[precision]
        self.storage = storage or precision
        self.acc = _NP[accumulate or precision]
        self.scaled = scaled
        self.coarse = coarse
        A = sps.csr_matrix(A, dtype=np.float64)
        self.n = A.shape[0]
        self.owner_ptr = np.asarray(owner_ptr, dtype=np.int64)
        self.sub_ptr = np.asarray(sub_ptr, dtype=np.int64)
        self.sub_idx = np.asarray(sub_idx, dtype=np.int64)
        n_s = self.owner_ptr.shape[0] - 1
        spc = np.ones(n_s, dtype=np.int64) if solvers_per_coarse is None else np.asarray(
            solvers_per_coarse, dtype=np.int64)
        scan = np.concatenate([[0], np.cumsum(spc)])
        self.coarse_ptr = self.owner_ptr[scan]
        self.n_c = spc.shape[0]
        # Local inverses: inverse in the preconditioner precision, then the storage cast.
        self.inv = []
        self.scale = []
        self.local_subset = None if local_subset is None else np.asarray(local_subset, np.int64)
        for i in (range(n_s) if local_subset is None else self.local_subset):
            idx = self.sub_idx[self.sub_ptr[i]:self.sub_ptr[i + 1]]
            if inverses is not None:
                self.inv.append((idx, np.asarray(inverses[i], dtype=np.float64).astype(self.acc)))
                self.scale.append(1.0)
                continue
            Ai = A[idx][:, idx].toarray().astype(self.pi)
            Ainv = np.linalg.inv(Ai).astype(np.float64)
            s = 1.0
            if scaled:
                s = float(np.abs(Ainv).max())
                s = float(np.float32(s))
            Aq = quantize(Ainv / s, self.storage)
            self.inv.append((idx, Aq.astype(self.acc)))
            self.scale.append(s)
        # Coarse restriction R_C (n_c x n), A_C = R_C A R_C^T, R0A = R_C A.
        cell = np.repeat(np.arange(self.n_c), np.diff(self.coarse_ptr))
        self.RC = sps.csr_matrix((np.ones(self.n), (cell, np.arange(self.n))),
                                 shape=(self.n_c, self.n))
        self.AC = (self.RC @ A @ self.RC.T).toarray().astype(self.pi)
        self.R0A = (self.RC @ A).tocsr().astype(self.pi)
        self.cho = sla.cho_factor(self.AC, lower=True) if coarse else None

    # -- pieces ---------------------------------------------------------------------------
    def local(self, x: np.ndarray) -> np.ndarray:
        """sum_i R_i^T s_i (A_i^-1 R_i x), ascending i; accumulation in `accumulate`."""
        y = np.zeros(self.n, dtype=self.pi)
        for (idx, Aq), s in zip(self.inv, self.scale):
            xi = x[idx].astype(self.acc)
            yi = (Aq @ xi) * self.acc(s)
            y[idx] += yi.astype(self.pi)
        return y

    def restrict(self, x: np.ndarray) -> np.ndarray:
        return np.add.reduceat(x.astype(self.pi), self.coarse_ptr[:-1]).astype(self.pi) \
            if self.n_c else np.zeros(0, self.pi)

    def prolong(self, yc: np.ndarray) -> np.ndarray:
        return np.repeat(yc, np.diff(self.coarse_ptr)).astype(self.pi)

    def coarse_solve(self, xc: np.ndarray) -> np.ndarray:
        return sla.cho_solve(self.cho, xc.astype(self.pi)).astype(self.pi)

    # -- operators ------------------------------------------------------------------------
    def apply(self, r: np.ndarray) -> np.ndarray:
        x = np.asarray(r).astype(self.pi)
        if self.kind == "additive":
            y = self.local(x)
            if self.coarse:
                y = y + self.prolong(self.coarse_solve(self.restrict(x)))
        else:
            y = self.local(x)
            z = (self.R0A @ y).astype(self.pi)
            y = y - self.prolong(self.coarse_solve(z))
        return y.astype(np.asarray(r).dtype)

    def T0(self, b: np.ndarray) -> np.ndarray:
        x = np.asarray(b).astype(self.pi)
        return self.prolong(self.coarse_solve(self.restrict(x))).astype(np.asarray(b).dtype)

    def dense(self) -> np.ndarray:
        """The preconditioner as an explicit n x n matrix (columns = apply(e_j))."""
        E = np.eye(self.n)
        return np.stack([self.apply(E[:, j]) for j in range(self.n)], axis=1)


class BlockSchwarzOracle:
    """FP64 Additive / Hybrid Schwarz for equal, contiguous, non-overlapping subdomains of k
    unknowns (the reference's layout), vectorised, with A_C factored by sparse LU."""

    def __init__(self, A: sps.csr_matrix, k: int, solvers_per_coarse=None,
                 kind: str = "additive"):
        A = sps.csr_matrix(A, dtype=np.float64)
        n = A.shape[0]
        if n % k:
            raise ValueError("the unknowns must split into subdomains of k")
        self.kind, self.n, self.k = kind, n, k
        n_s = n // k
        self.n_s = n_s
        rows = np.repeat(np.arange(n), np.diff(A.indptr))
        cols = A.indices.astype(np.int64)
        own = (rows // k) == (cols // k)
        D = np.zeros((n_s, k, k))
        D[rows[own] // k, rows[own] % k, cols[own] % k] = A.data[own]
        self.inv = np.linalg.inv(D)                 # A_i^-1 (solvers.py:351)
        spc = np.ones(n_s, np.int64) if solvers_per_coarse is None else np.asarray(
            solvers_per_coarse, np.int64)
        self.coarse_ptr = k * np.concatenate([[0], np.cumsum(spc)])
        n_c = spc.shape[0]
        cell = np.repeat(np.arange(n_c), np.diff(self.coarse_ptr))
        RC = sps.csr_matrix((np.ones(n), (cell, np.arange(n))), shape=(n_c, n))
        self.R0A = (RC @ A).tocsr()                 # R_C A (solvers.py:890-906)
        self.lu = spla.splu((self.R0A @ RC.T).tocsc())   # A_C = R_C A R_C^T

    def local(self, x: np.ndarray) -> np.ndarray:
        return np.matmul(self.inv, x.reshape(self.n_s, self.k, 1)).reshape(-1)

    def restrict(self, x: np.ndarray) -> np.ndarray:
        return np.add.reduceat(x, self.coarse_ptr[:-1])

    def prolong(self, yc: np.ndarray) -> np.ndarray:
        return np.repeat(yc, np.diff(self.coarse_ptr))

    def coarse_solve(self, xc: np.ndarray) -> np.ndarray:
        return self.lu.solve(xc)

    def apply(self, r: np.ndarray) -> np.ndarray:
        r = np.asarray(r, dtype=np.float64)
        y = self.local(r)
        if self.kind == "additive":
            return y + self.prolong(self.coarse_solve(self.restrict(r)))
        return y - self.prolong(self.coarse_solve(self.R0A @ y))

    def T0(self, b: np.ndarray) -> np.ndarray:
        return self.prolong(self.coarse_solve(self.restrict(np.asarray(b, np.float64))))
