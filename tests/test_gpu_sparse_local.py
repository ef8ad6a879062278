"""The sparse local path: `SparseBatchSolver`s in the Schwarz preconditioners.

The reference hands a sparse local solver the A_i as a CSRBatch
(src/dd_solvers/dd_solvers/solvers.py:610-616, 660-719) and factorises them with cuDSS's
batch API (`BatchCUDSS`, :390-403). Here `BatchSparseCholesky` fills that slot with the
nested-dissection sweeps. Its local solves are exact up to rounding, so the bar is the
oracle with FP64 local inverses: iterates within 1e-10 relative, iterations within +-1.
A user-written SparseBatchSolver runs through the generic path on the reference's
layout, which is checked field by field against the oracle's restatement.
"""

import copy

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla
import torch

from helpers import dg24, dg24_permuted
from oracle.kernels import local_csr_batch as ref_csr_batch
from oracle.pcg import cg as oracle_cg
from oracle.schwarz import SchwarzOracle

import dd_solvers as dds
from dd_solvers import problems as P
from dd_solvers.solvers import CSRBatch

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _torch_csr(A, dtype=torch.float64):
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                   torch.as_tensor(A.data), A.shape).to(device=dev, dtype=dtype)


def _parity(pb, kind, local, rtol=1e-8, use_graph=False, check_every=8, coarse=None,
            it_tol=1, x_tol=1e-10):
    d = pb.decomposition
    O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind=kind,
                      storage="float64")
    xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply,
                       T0=O.T0 if kind == "hybrid" else None, rtol=rtol)
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    s = dds.CG(cls(torch.float64, local, coarse or dds.DenseCholesky()), rtol=rtol,
               bsr_matmul=False, check_every=check_every, use_graph=use_graph)
    s.setup(_torch_csr(pb.A), decomposition=d)
    x, m = s.solve(torch.as_tensor(pb.b, device=dev))
    assert abs(m["iterations"] - mo["iterations"]) <= it_tol, (m["iterations"], mo["iterations"])
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    assert err <= x_tol, err
    res = np.linalg.norm(pb.b - pb.A @ x.cpu().numpy()) / np.linalg.norm(pb.b)
    assert res < rtol * 1.01
    return x, m, s


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_q1_3d_overlap1(kind):
    """Uneven, overlapping 3D boxes: the forward sweep gathers through the index lists."""
    pb = P.make_problem("q1_3d", 24, box=8, overlap=1, seed=1, lognormal=True)
    _, _, s = _parity(pb, kind, dds.BatchSparseCholesky())
    pc = s.preconditioner
    assert pc._device_apply and pc._sparse_local and pc.bytes_per_apply > 0


def test_p1_2d_overlap2():
    pb = P.make_problem("p1_2d", 128, box=8, overlap=2, seed=0)
    _parity(pb, "hybrid", dds.BatchSparseCholesky(leaf_size=32))


def test_kmeans_uneven():
    pb = P.make_problem("p1_2d_kmeans", 96, overlap=2, seed=5, n_clusters=12, jitter=0.2)
    _parity(pb, "hybrid", dds.BatchSparseCholesky())


def test_sparse_coarse_and_sparse_local():
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    _parity(pb, "hybrid", dds.BatchSparseCholesky(leaf_size=16),
            coarse=dds.SparseCholesky(leaf_size=8, top_max=16))


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_fp32_factor_fp64_vectors(kind):
    """The factor stored in FP32 (computed in FP64), vectors and sums in FP64: the precision
    form of the sparse local slot. Bar as for the FP32 inverses: iterations within 2 of the
    FP64 oracle's, solution within 1e-6, true residual below rtol."""
    pb = P.make_problem("q1_3d", 24, box=8, overlap=1, seed=1, lognormal=True)
    _, _, s = _parity(pb, kind, dds.BatchSparseCholesky(precision=torch.float32), it_tol=2,
                      x_tol=1e-6)
    ls = s.preconditioner.local_solver
    assert ls._inner.fac.dtype == torch.float32 and ls._inner.dtype == torch.float64


def test_graph_replay_matches_eager():
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    x1, m1, _ = _parity(pb, "hybrid", dds.BatchSparseCholesky(leaf_size=16), check_every=4)
    x2, m2, _ = _parity(pb, "hybrid", dds.BatchSparseCholesky(leaf_size=16), check_every=4,
                        use_graph=True)
    assert m1["iterations"] == m2["iterations"]
    assert torch.equal(x1, x2)


# ---------------------------------------------------------------------------------------
# the reference's fixture, and the generic path for user solvers
# ---------------------------------------------------------------------------------------

def _dg24_setup(solver, problem_dtype=torch.float64):
    A, b, f2s, s2c, k = dg24()
    matrix = _torch_csr(A, problem_dtype)
    rhs = torch.as_tensor(b, dtype=problem_dtype, device=dev)
    fine_to_dofs = torch.arange(24, device=dev, dtype=torch.int32).reshape(8, 3)
    s2c_t = torch.as_tensor(s2c)
    n_coarse = int(s2c_t.max()) + 1
    perm = dds.asm_permutation(torch.as_tensor(f2s), s2c_t, fine_to_dofs).to(dev)
    inv_perm = dds.inverse_permutation(perm)
    matrix = dds.permute_csr_tensor(matrix, inv_perm)
    solver.setup(matrix=matrix, dofs_per_solver=k,
                 solvers_per_coarse=s2c_t.bincount(minlength=n_coarse).to(dev),
                 block_size=fine_to_dofs.shape[1])
    x, _ = solver.solve(rhs[perm])
    solver.destroy()
    xr = x[inv_perm].cpu().numpy().astype(np.float64)
    return np.linalg.norm(sps.csr_matrix(A) @ xr - b)


DG24 = [
    (dds.CG(dds.AdditiveSchwarz(torch.float64, dds.BatchSparseCholesky(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.BatchSparseCholesky(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.BatchSparseCholesky(), dds.SparseCholesky())), torch.float64),
    # an FP32 Hybrid preconditioner stalls CG near 1e-7; the reference tests it at 1e-6
    (dds.CG(dds.HybridSchwarz(torch.float32, dds.BatchSparseCholesky(), dds.DenseCholesky()), rtol=1e-6), torch.float64),
    (dds.CG(dds.AdditiveSchwarz(torch.float32, dds.BatchSparseCholesky(), dds.DenseCholesky())), torch.float32),
]


@pytest.mark.parametrize("solver,problem_dtype", DG24, ids=[f"{s}-{d}" for s, d in DG24])
def test_dg24_residual(solver, problem_dtype):
    """src/dd_solvers/tests/test_solvers.py:149-176 with the sparse local slot."""
    assert _dg24_setup(copy.deepcopy(solver), problem_dtype) < 1e-10


class DenseFromCSR(dds.SparseBatchSolver):
    """A user-written sparse local solver: keeps the CSRBatch it is given and solves with
    dense inverses built from it (torch only)."""

    def setup(self, matrix_batch: CSRBatch):
        self.batch = CSRBatch(*(t.clone() for t in matrix_batch))
        n_s, k = int(matrix_batch.nrows.shape[0]), int(matrix_batch.nrows[0])
        ro = matrix_batch.row_offsets.to(torch.int64).view(n_s, k + 1)
        counts = (ro[:, 1:] - ro[:, :-1]).reshape(-1)
        row = torch.repeat_interleave(torch.arange(n_s * k, device=dev), counts)
        M = torch.zeros(n_s * k * k, dtype=matrix_batch.values.dtype, device=dev)
        M[row * k + matrix_batch.col_indices.to(torch.int64)] = matrix_batch.values
        self.inv = torch.linalg.inv(M.view(n_s, k, k))

    def solve(self, rhs_batch):
        return torch.bmm(self.inv, rhs_batch.reshape(self.inv.shape[0], -1, 1)).reshape(-1)


def test_user_sparse_batch_solver_gets_reference_layout():
    A, b, perm, Ap, k, spc = dg24_permuted()
    user = DenseFromCSR()
    s = dds.CG(dds.HybridSchwarz(torch.float64, user, dds.DenseCholesky()), rtol=1e-12)
    s.setup(_torch_csr(Ap), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
            block_size=3)
    want = ref_csr_batch(Ap.indptr, Ap.indices, Ap.data, k)
    for got, w in zip(user.batch, want):
        assert np.array_equal(got.cpu().numpy(), w)
    assert not s.preconditioner._device_apply
    x, m = s.solve(torch.as_tensor(b[perm], device=dev))
    n = Ap.shape[0]
    ptr = np.arange(0, n + 1, k)
    O = SchwarzOracle(Ap, ptr, ptr, np.arange(n), spc, kind="hybrid")
    xo, mo = oracle_cg(lambda v: Ap @ v, b[perm], precond=O.apply, T0=O.T0, rtol=1e-12)
    assert abs(m["iterations"] - mo["iterations"]) <= 1
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)


def test_user_sparse_batch_solver_needs_reference_subdomains():
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    s = dds.CG(dds.HybridSchwarz(torch.float64, DenseFromCSR(), dds.DenseCholesky()),
               bsr_matmul=False)
    with pytest.raises(ValueError, match="equal, non-overlapping"):
        s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)


@pytest.mark.parametrize("dtype,precision,tol", [(torch.float64, None, 1e-12),
                                                   (torch.float32, None, 2e-5),
                                                   (torch.float64, torch.float32, 2e-6)])
def test_batch_solver_standalone(dtype, precision, tol):
    """setup(CSRBatch) / solve((n_s, k)) -> flat, as BatchCUDSS: against SuperLU per block."""
    rng = np.random.default_rng(7)
    n_s, k = 9, 40
    blocks = []
    for _ in range(n_s):
        B = sps.random(k, k, density=0.1, random_state=rng)
        blocks.append((B + B.T + sps.diags(np.full(k, 6.0))).tocsr())
    Ab = sps.block_diag(blocks).tocsr()
    Ab.sort_indices()
    fields = ref_csr_batch(Ab.indptr, Ab.indices, Ab.data, k)
    batch = CSRBatch(*(torch.as_tensor(f, device=dev) for f in fields))
    batch = batch._replace(values=batch.values.to(dtype))
    solver = dds.BatchSparseCholesky(leaf_size=8, precision=precision)
    solver.setup(batch)
    rhs = rng.standard_normal((n_s, k))
    x = solver.solve(torch.as_tensor(rhs, dtype=dtype, device=dev))
    assert x.shape == (n_s * k,) and x.dtype == dtype
    want = np.concatenate([spla.spsolve(Bk.tocsc(), r) for Bk, r in zip(blocks, rhs)])
    assert np.linalg.norm(x.double().cpu().numpy() - want) <= tol * np.linalg.norm(want)


def test_factor_precision_checks():
    with pytest.raises(ValueError, match="unsupported factor precision"):
        dds.BatchSparseCholesky(precision=torch.float16)
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    s = dds.CG(dds.HybridSchwarz(torch.float32, dds.BatchSparseCholesky(precision=torch.float64),
                                 dds.DenseCholesky()), bsr_matmul=False)
    with pytest.raises(ValueError, match="FP64 factor needs FP64 values"):
        s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
