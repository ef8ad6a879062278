"""End-to-end GPU parity of CG + Schwarz against the CPU oracle.

Bars: FP64 iterates within 1e-10 relative of the oracle's, iteration counts within +-1,
and the reference's own residual check on its fixture (||A x - b|| < 1e-10).
"""

import copy

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from helpers import dg24, dg24_permuted
from oracle.pcg import cg as oracle_cg
from oracle.schwarz import SchwarzOracle

import dd_solvers as dds
from dd_solvers import problems as P

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _torch_csr(A, dtype=torch.float64):
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                   torch.as_tensor(A.data), A.shape).to(device=dev, dtype=dtype)


# ---------------------------------------------------------------------------------------
# the reference's fixture and configurations (src/dd_solvers/tests/test_solvers.py)
# ---------------------------------------------------------------------------------------

CONFIGS = [
    (dds.CG(bsr_matmul=False), torch.float32),
    (dds.CG(bsr_matmul=True), torch.float64),
    (dds.CG(dds.AdditiveSchwarz(torch.float64, dds.Inv(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.AdditiveSchwarz(torch.float32, dds.Inv(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.AdditiveSchwarz(torch.float32, dds.Inv(), dds.DenseCholesky())), torch.float32),
    (dds.CG(dds.AdditiveSchwarz(torch.float32, dds.Inv(torch.float16), dds.DenseCholesky())), torch.float32),
    (dds.CG(dds.AdditiveSchwarz(torch.float64, dds.LU(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.AdditiveSchwarz(torch.float64, dds.Cholesky(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.float32), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.float16), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.bfloat16), dds.DenseCholesky())), torch.float64),
    (dds.CG(dds.HybridSchwarz(torch.float32, dds.Inv(), dds.DenseCholesky()), rtol=1e-6), torch.float32),
    (dds.CG(dds.HybridSchwarz(torch.float32, dds.Inv(), dds.DenseCholesky()), rtol=1e-6), torch.float64),
]


@pytest.mark.parametrize("solver,problem_dtype", CONFIGS, ids=[f"{s}-{d}" for s, d in CONFIGS])
def test_sparse_solver_dg24(solver, problem_dtype):
    """src/dd_solvers/tests/test_solvers.py:149-176 with DenseCholesky in the coarse slot."""
    solver = copy.deepcopy(solver)
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
    # the reference's check, evaluated in FP64 on the host
    xr = x[inv_perm].cpu().numpy().astype(np.float64)
    Ah = sps.csr_matrix(A)
    assert np.linalg.norm(Ah @ xr - b) < 1e-10


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_dg24_iterates_match_oracle(kind):
    A, b, perm, Ap, k, spc = dg24_permuted()
    n = Ap.shape[0]
    ptr = np.arange(0, n + 1, k)
    O = SchwarzOracle(Ap, ptr, ptr, np.arange(n), spc, kind=kind)
    xo, mo = oracle_cg(lambda v: Ap @ v, b[perm], precond=O.apply,
                       T0=O.T0 if kind == "hybrid" else None, rtol=1e-12)
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    s = dds.CG(cls(torch.float64, dds.Inv(), dds.DenseCholesky()), rtol=1e-12)
    s.setup(_torch_csr(Ap), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
            block_size=3)
    x, m = s.solve(torch.as_tensor(b[perm], device=dev))
    assert abs(m["iterations"] - mo["iterations"]) <= 1
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    k_ = min(len(m["residual norms"]), len(mo["residual norms"]))
    assert np.allclose(m["residual norms"][:k_], mo["residual norms"][:k_], rtol=1e-8, atol=1e-30)


def test_cg_condition_number_estimate():
    # src/dd_solvers/tests/test_solvers.py:179-184, on the GPU
    A = torch.diag(torch.tensor([0.5, 4.0, 21.0], dtype=torch.float64)).to_sparse_csr().to(dev)
    s = dds.CG(preconditioner=None, rtol=1e-9, estimate_cond=True, bsr_matmul=False)
    s.setup(A)
    _, meta = s.solve(torch.ones(3, dtype=torch.float64, device=dev))
    assert np.isclose(meta["condition number estimate"], 42.0)


def test_convergence_not_achieved():
    A, b, perm, Ap, k, spc = dg24_permuted()
    s = dds.CG(dds.AdditiveSchwarz(torch.float64, dds.Inv(), dds.DenseCholesky()), maxiter=1,
               rtol=1e-12)
    s.setup(_torch_csr(Ap), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
            block_size=3)
    with pytest.raises(dds.ConvergenceNotAchieved) as e:
        s.solve(torch.as_tensor(b[perm], device=dev))
    assert e.value.metadata["iterations"] == 0
    assert len(e.value.metadata["residual norms"]) == 1


def test_missing_block_size():
    A, b, perm, Ap, k, spc = dg24_permuted()
    with pytest.raises(ValueError):
        dds.CG().setup(_torch_csr(Ap))


# ---------------------------------------------------------------------------------------
# model problems: overlap, uneven subdomains, lognormal coefficients
# ---------------------------------------------------------------------------------------

def _parity(pb, kind, storage="float64", rtol=1e-8, check_every=8, use_graph=False,
            bsr=None, it_tol=1, x_tol=1e-10, layout="sym"):
    d = pb.decomposition
    O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind=kind,
                      storage=storage)
    xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply,
                       T0=O.T0 if kind == "hybrid" else None, rtol=rtol)
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    tdt = {"float64": torch.float64, "float32": torch.float32, "float16": torch.float16,
           "bfloat16": torch.bfloat16}[storage]
    s = dds.CG(cls(torch.float64, dds.Inv(tdt, layout=layout), dds.DenseCholesky()), rtol=rtol,
               bsr_matmul=bsr is not None, check_every=check_every, use_graph=use_graph)
    kw = {"block_size": bsr} if bsr else {}
    s.setup(_torch_csr(pb.A), decomposition=d, **kw)
    x, m = s.solve(torch.as_tensor(pb.b, device=dev))
    assert abs(m["iterations"] - mo["iterations"]) <= it_tol, (m["iterations"], mo["iterations"])
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    assert err <= x_tol, err
    res = np.linalg.norm(pb.b - pb.A @ x.cpu().numpy()) / np.linalg.norm(pb.b)
    assert res < rtol * 1.01
    return x, m


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_c1_p1_2d_overlap2(kind):
    """2D P1 on 128^2, 8x8-element boxes, overlap 2, FP64, rtol 1e-8 (the CPU-runnable
    configuration), bit-for-bit-level agreement with the oracle."""
    pb = P.make_problem("p1_2d", 128, box=8, overlap=2, seed=0)
    _parity(pb, kind)


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_q1_3d_lognormal_overlap1(kind):
    pb = P.make_problem("q1_3d", 24, box=8, overlap=1, seed=1, lognormal=True)
    _parity(pb, kind)


def test_q1_3d_full_layout_matches_oracle():
    pb = P.make_problem("q1_3d", 24, box=8, overlap=1, seed=1, lognormal=True)
    _parity(pb, "hybrid", layout="full")


@pytest.mark.parametrize("b", [2, 4, 8])
def test_q1_3d_bsr_block_sizes(b):
    # n = 24^3: every b divides n, no padding
    pb = P.make_problem("q1_3d", 25, box=8, overlap=1, seed=3, lognormal=True)
    assert pb.A.shape[0] % b == 0
    _parity(pb, "hybrid", bsr=b)


@pytest.mark.parametrize("b", [2, 4, 8])
def test_q1_3d_bsr_padded_tail(b):
    """n = 23^3 = 12167, which none of b = 2, 4, 8 divides (as 63^3 at 64^3): the operator is
    padded to a whole number of block rows and the PCG vectors carry a zero tail. Iterates,
    iteration count and the residual history against the oracle."""
    pb = P.make_problem("q1_3d", 24, box=8, overlap=1, seed=3, lognormal=True)
    n = pb.A.shape[0]
    assert n % b != 0
    d = pb.decomposition
    O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid")
    xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply, T0=O.T0, rtol=1e-8)
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(), dds.DenseCholesky()), rtol=1e-8)
    s.setup(_torch_csr(pb.A), decomposition=d, block_size=b)
    assert s.matrix.shape[0] == -(-n // b) * b > n
    x, m = s.solve(torch.as_tensor(pb.b, device=dev))
    assert x.shape[0] == n
    assert abs(m["iterations"] - mo["iterations"]) <= 1
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    k_ = min(len(m["residual norms"]), len(mo["residual norms"]))
    assert np.allclose(m["residual norms"][:k_], mo["residual norms"][:k_], rtol=1e-8, atol=0)
    # the padded tail of the working vectors stays zero
    assert not s._buf.x[n:].any() and not s._buf.r[n:].any()


@pytest.mark.parametrize("use_graph", [False, True])
def test_setup_again_with_larger_problem(use_graph):
    """One CG + HybridSchwarz object set up on a small problem, solved, then set up on a
    larger one (larger n and n_c) and solved again, single right-hand side and blocks of
    16 (full-inverse MFMA store, then the symmetric matrix-core tiles): every buffer and
    captured graph follows the new sizes (results against the oracle)."""
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.float16, scaled=True, layout="mfma"),
                                 dds.DenseCholesky()), rtol=1e-8, bsr_matmul=False,
               check_every=4, use_graph=use_graph)
    for N, box in ((10, 3), (16, 4)):
        pb = P.make_problem("q1_3d", N, box=box, overlap=1, seed=N, lognormal=True)
        d = pb.decomposition
        s.setup(_torch_csr(pb.A), decomposition=d)
        O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid",
                          inverses=s.preconditioner.stored_local_inverses())
        xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply, T0=O.T0, rtol=1e-8)
        x, m = s.solve(torch.as_tensor(pb.b, device=dev))
        assert abs(m["iterations"] - mo["iterations"]) <= 2
        assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
        B = np.random.default_rng(N).uniform(-1, 1, (pb.A.shape[0], 16))
        X, mb = s.solve(torch.as_tensor(B, device=dev))
        res = np.linalg.norm(B - pb.A @ X.cpu().numpy(), axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() < 1.01e-8
    # the symmetric matrix-core tiles (one vector and blocks of 16 on the same store)
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.float16, scaled=True, layout="sym_mfma",
                                                        accumulate=torch.float32),
                                 dds.DenseCholesky()), rtol=1e-8, bsr_matmul=False,
               check_every=4, use_graph=use_graph)
    for N, box in ((10, 3), (16, 4)):
        pb = P.make_problem("q1_3d", N, box=box, overlap=1, seed=N, lognormal=True)
        d = pb.decomposition
        s.setup(_torch_csr(pb.A), decomposition=d)
        assert s.preconditioner._msym is not None
        O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid",
                          inverses=s.preconditioner.stored_local_inverses())
        xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply, T0=O.T0, rtol=1e-8)
        x, m = s.solve(torch.as_tensor(pb.b, device=dev))
        assert abs(m["iterations"] - mo["iterations"]) <= 2
        assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
        B = np.random.default_rng(N).uniform(-1, 1, (pb.A.shape[0], 16))
        X, mb = s.solve(torch.as_tensor(B, device=dev))
        res = np.linalg.norm(B - pb.A @ X.cpu().numpy(), axis=0) / np.linalg.norm(B, axis=0)
        assert res.max() < 1.01e-8
    # and the plain FP64 preconditioner on a new object graph, set up twice the same way
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(), dds.DenseCholesky()), rtol=1e-8,
               bsr_matmul=False, check_every=4, use_graph=use_graph)
    for N, box in ((10, 3), (16, 4)):
        pb = P.make_problem("q1_3d", N, box=box, overlap=1, seed=N, lognormal=True)
        d = pb.decomposition
        s.setup(_torch_csr(pb.A), decomposition=d)
        O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid")
        xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply, T0=O.T0, rtol=1e-8)
        x, m = s.solve(torch.as_tensor(pb.b, device=dev))
        assert abs(m["iterations"] - mo["iterations"]) <= 1
        assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("check_every", [4, 8, 3])
def test_graph_replay_matches_eager(check_every):
    """Captured blocks (with the short blocks near convergence when check_every >= 4) give
    the eager solve's iterations, residual norms and bits."""
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    x1, m1 = _parity(pb, "hybrid", check_every=check_every)
    x2, m2 = _parity(pb, "hybrid", check_every=check_every, use_graph=True)
    assert m1["iterations"] == m2["iterations"]
    assert m1["residual norms"] == m2["residual norms"]
    assert torch.equal(x1, x2)


@pytest.mark.parametrize("storage", ["float32", "float16", "bfloat16"])
def test_low_precision_inverses(storage):
    pb = P.make_problem("q1_3d", 16, box=4, overlap=1, seed=4, lognormal=True)
    # iterations may differ more with inverses rounded independently on each side; the
    # solution still meets rtol
    _parity(pb, "hybrid", storage=storage, it_tol=2, x_tol=1e-6)


@pytest.mark.parametrize("storage,accumulate,layout", [
    (torch.float32, None, "sym"), (torch.float16, None, "sym"), (torch.bfloat16, None, "sym"),
    (torch.float16, torch.float32, "sym_mfma"), (torch.bfloat16, torch.float32, "sym_mfma")])
def test_low_precision_against_stored_inverses(storage, accumulate, layout):
    """The oracle applies exactly the inverses the device stores (rounding and scale
    included). With FP64 accumulation the two preconditioners differ by rounding only, so
    the FP64 bar holds (iterates 1e-10, iterations +-1); the matrix-core apply keeps ~22-24
    bits of x and accumulates in FP32 (iterates 1e-6, iterations +-1)."""
    pb = P.make_problem("q1_3d", 18, box=8, overlap=1, seed=4, lognormal=True)
    d = pb.decomposition
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(storage, scaled=storage == torch.float16,
                                                        accumulate=accumulate, layout=layout),
                                 dds.DenseCholesky()), rtol=1e-8, bsr_matmul=False)
    s.setup(_torch_csr(pb.A), decomposition=d)
    assert s.preconditioner.local.layout == layout
    O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind="hybrid",
                      inverses=s.preconditioner.stored_local_inverses())
    xo, mo = oracle_cg(lambda v: pb.A @ v, pb.b, precond=O.apply, T0=O.T0, rtol=1e-8)
    x, m = s.solve(torch.as_tensor(pb.b, device=dev))
    assert abs(m["iterations"] - mo["iterations"]) <= 1
    err = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    assert err <= (1e-6 if layout == "sym_mfma" else 1e-10), err
