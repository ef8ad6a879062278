"""Differentiable Schwarz-preconditioned solve: `torch.autograd.Function` around
`CG.solve` for the Q1 Poisson problem -div(exp(theta) grad u) = f on an N^3 grid.

The reference solves with fixed matrices and has no autograd; this is a new capability on
top of the drop-in path. Forward: assemble A(theta) on the GPU (A.data = G exp(theta) with
the host-built assembly operator G, through the CSR kernel), set up the Schwarz
preconditioner, solve. Backward: A is symmetric, so the adjoint solve A lam = dL/du runs the
same CG with the same preconditioner, and
    dL/dtheta_e = -exp(theta_e) h lam_e^T Khat u_e      (dds_q1_element_grad)
    dL/df       = lam.

    op = Q1DiffSolve(N=48, box=8, overlap=1, storage=torch.float32, rtol=1e-12)
    u = op(theta, f)                 # theta: (N^3,) float64 cuda, f: ((N-1)^3,) float64 cuda
    loss = 0.5 * ((u - u_obs) ** 2).sum(); loss.backward()
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import problems as P
from .solvers import CG, DenseCholesky, HybridSchwarz, Inv, _CSROperator

__all__ = ["Q1DiffSolve"]


class Q1DiffSolve:
    """Callable (theta, f) -> u with gradients for theta and f."""

    def __init__(self, N: int, box: int = 8, overlap: int = 1,
                 storage: torch.dtype = torch.float32, rtol: float = 1e-12,
                 maxiter: int = 5000, device="cuda"):
        self.N = N
        self.device = dev = torch.device(device)
        self.rtol = rtol
        self.maxiter = maxiter
        self.storage = storage
        m = N - 1
        self.n = m ** 3
        dec = P.box_decomposition(m, 3, box, overlap)
        self.decomposition = dec
        # pattern of A in the subdomain numbering, and for each permuted entry the entry of
        # the natural-order pattern it comes from
        A0 = P.assemble_q1_3d(N)
        idx = A0.copy()
        idx.data = np.arange(A0.nnz, dtype=np.float64)
        Ap = idx[dec.perm][:, dec.perm].tocsr()
        Ap.sort_indices()
        src = Ap.data.astype(np.int64)
        G = P.q1_assembly_operator(N)[src]  # rows in permuted-entry order
        self._crow = torch.as_tensor(Ap.indptr.astype(np.int64), device=dev)
        self._col = torch.as_tensor(Ap.indices.astype(np.int32), device=dev)
        Gt = torch.sparse_csr_tensor(torch.as_tensor(G.indptr.astype(np.int64)),
                                     torch.as_tensor(G.indices.astype(np.int32)),
                                     torch.as_tensor(G.data), G.shape).to(dev)
        self._assembly = _CSROperator(Gt)
        perm = torch.as_tensor(dec.perm, device=dev)
        self.perm = perm
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(perm.shape[0], device=dev)
        self.inv_perm = inv
        # natural interior node -> storage position, for the element-gradient kernel
        self._map = inv.to(torch.int32)
        self.solver = None

    def matrix(self, theta: torch.Tensor) -> torch.Tensor:
        """A(theta) in the subdomain numbering, as a torch sparse CSR on the device."""
        k = torch.exp(theta.detach().to(torch.float64)).contiguous()
        vals = self._assembly @ k
        return torch.sparse_csr_tensor(self._crow, self._col, vals, (self.n, self.n))

    def setup(self, theta: torch.Tensor) -> None:
        s = CG(HybridSchwarz(torch.float64, Inv(self.storage), DenseCholesky()), rtol=self.rtol,
               maxiter=self.maxiter, bsr_matmul=False)
        s.setup(self.matrix(theta), decomposition=self.decomposition)
        self.solver = s

    def _solve_natural(self, b: torch.Tensor, solver: CG | None = None) -> tuple[torch.Tensor, dict]:
        solver = solver or self.solver
        x, meta = solver.solve(b.to(torch.float64).contiguous()[self.perm].contiguous())
        return x[self.inv_perm], meta

    def element_gradient(self, theta, lam, u) -> torch.Tensor:
        """-exp(theta_e) h lam_e^T Khat u_e; lam, u in the subdomain numbering."""
        grad = torch.empty(self.N ** 3, dtype=torch.float64, device=self.device)
        theta = theta.detach().to(torch.float64).contiguous()
        with torch.cuda.device(self.device):
            _lib.call("dds_q1_element_grad", self.N, theta.data_ptr(), lam.data_ptr(),
                      u.data_ptr(), self._map.data_ptr(), grad.data_ptr(), _lib.stream())
        return grad

    def __call__(self, theta: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
        return _Q1Solve.apply(theta, f, self)


class _Q1Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, f, op: Q1DiffSolve):
        _lib.require_device(theta, f)
        op.setup(theta)
        u, meta = op._solve_natural(f)
        ctx.op = op
        # the solver of this theta: a later forward call replaces op.solver, and this node's
        # adjoint must still be solved with A(theta) of its own forward
        ctx.solver = op.solver
        ctx.forward_iterations = meta["iterations"]
        ctx.save_for_backward(theta.detach(), u)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        op = ctx.op
        theta, u = ctx.saved_tensors
        # adjoint solve with the same operator and preconditioner (A is symmetric)
        lam, meta = op._solve_natural(grad_u, ctx.solver)
        ctx.solver = None
        op.last_adjoint_iterations = meta["iterations"]
        g_theta = op.element_gradient(theta, lam[op.perm].contiguous(), u[op.perm].contiguous())
        return g_theta.to(theta.dtype), lam.to(grad_u.dtype), None
