"""The sparse coarse solver (SparseCholesky: dds_nd_forward / dds_nd_backward) on the GPU:
against scipy's sparse direct solve and the host restatement of its sweeps, and inside the
Schwarz preconditioners on the reference's DG experiments, including the published
S10 / S8 / S8 runs whose 131,072-cell coarse space is beyond a dense factor."""

import json
import os

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla
import torch

import dd_solvers as dds
from dd_solvers import dg
from dd_solvers import sparse_coarse as sc

from oracle import nd_plan as onp
from test_sparse_coarse import dg_coarse, lap2d, random_spd_pieces

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _torch_csr(A, dtype=torch.float64):
    A = sps.csr_matrix(A)
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr.astype(np.int64)),
                                   torch.as_tensor(A.indices.astype(np.int64)),
                                   torch.as_tensor(A.data), A.shape).to(dev).to(dtype)


CASES = {
    "one": (lambda: lap2d(1, 1), 32, 4096),
    "lap_top": (lambda: lap2d(40, 33), 16, 300),
    "lap_notop": (lambda: lap2d(40, 33), 16, 0),
    "pieces": (lambda: random_spd_pieces(), 8, 16),
    "dg_s6_s4": (lambda: dg_coarse(6, "S4"), 8, 64),
    "lap_big": (lambda: lap2d(300, 257), 32, 4096),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_solve_matches_superlu(name, dtype):
    mk, leaf, top = CASES[name]
    A = mk()
    print(f"SparseCholesky {name} n={A.shape[0]} {dtype}", flush=True)
    S = dds.SparseCholesky(leaf_size=leaf, top_max=top)
    S.setup(_torch_csr(A, dtype))
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    x, md = S.solve(torch.as_tensor(b, device=dev).to(dtype))
    x = x.double().cpu().numpy()
    assert md == {}
    xr = spla.spsolve(A.tocsc(), b)
    tol = 1e-11 if dtype == torch.float64 else 2e-4
    assert np.linalg.norm(x - xr) <= tol * np.linalg.norm(xr)
    if dtype == torch.float64 and A.shape[0] < 5000:
        # the device factor (Gauss-Jordan sweeps of the fronts, dense Cholesky of the top)
        # against the restated one (Cholesky of every front), and the host restatement of
        # the sweeps over the same tables and factor
        fac = S.fac.cpu().numpy()
        fh = onp.factorize_host(S.plan, S.plan.Ap.data)
        assert np.abs(fac - fh).max() <= 1e-12 * np.abs(fh).max()
        xh = onp.solve_host(S.plan, fac, b)
        assert np.abs(x - xh).max() <= 1e-13 * np.abs(xh).max()


def test_solve_is_deterministic_and_reusable():
    A = lap2d(120, 97)
    S = dds.SparseCholesky(leaf_size=16, top_max=512)
    S.setup(_torch_csr(A))
    rng = np.random.default_rng(0)
    b1 = torch.as_tensor(rng.standard_normal(A.shape[0]), device=dev)
    b2 = torch.as_tensor(rng.standard_normal(A.shape[0]), device=dev)
    x1, _ = S.solve(b1)
    x2, _ = S.solve(b2)
    x1b, _ = S.solve(b1)
    assert torch.equal(x1, x1b)
    assert not torch.equal(x1, x2)


def _dg_solver(kind, coarse, storage=None, pi=torch.float64):
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    return dds.CG(cls(pi, dds.Inv(storage), coarse), rtol=1e-9, maxiter=2000)


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_schwarz_with_sparse_coarse_matches_dense_coarse(kind):
    """DG p = 1 on S6 with S4 subdomains and S4 coarse cells (512): the same iterations as
    the dense coarse factor, solutions to rounding."""
    print(f"DG S6 S4 S4 {kind}: sparse vs dense coarse", flush=True)
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(6), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S4", "S4")
    out = {}
    for name, coarse in (("dense", dds.DenseCholesky()),
                         ("sparse", dds.SparseCholesky(leaf_size=8, top_max=64))):
        s = _dg_solver(kind, coarse)
        s.setup(_torch_csr(A), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
                block_size=3)
        x, md = s.solve(torch.as_tensor(b, device=dev))
        out[name] = (x.cpu().numpy(), md)
    (xd, md_d), (xs, md_s) = out["dense"], out["sparse"]
    # the two coarse solves differ by rounding only; CG carries that into the late iterates
    assert abs(md_d["iterations"] - md_s["iterations"]) <= 1
    assert np.allclose(md_d["residual norms"][:10], md_s["residual norms"][:10], rtol=1e-10, atol=0)
    assert np.linalg.norm(xs - xd) <= 1e-8 * np.linalg.norm(xd)


PUB = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dg_published.json")))


@pytest.fixture(scope="module")
def s10_s8_s8():
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(10), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S8", "S8")
    return A, b, k, spc


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_published_s10_s8_s8_runs_on_the_gpu(s10_s8_s8, kind):
    """The reference's CG(Additive / HybridSchwarz(float64, Inv(), CUDSS)) runs, p = 1, S10
    (6,291,456 DoFs), S8 subdomains and coarse cells (131,072 of each), rtol 1e-9, with the
    sparse coarse solver in the CUDSS slot: the published iteration count and residual
    history (results/experiment_precisions.csv)."""
    A, b, k, spc = s10_s8_s8
    ref = PUB["runs"][f"p1_S10_S8_S8_{kind}"]
    print(f"DG p1 S10 S8 S8 {kind} (published {ref['iterations']} it)", flush=True)
    s = _dg_solver(kind, dds.SparseCholesky())
    s.setup(_torch_csr(A), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
            block_size=3)
    x, md = s.solve(torch.as_tensor(b, device=dev))
    x = x.cpu().numpy()
    assert md["iterations"] == ref["iterations"]
    got, pub = np.array(md["residual norms"]), np.array(ref["residual_norms"])
    assert np.allclose(got[:10], pub[:10], rtol=1e-10, atol=0)
    assert np.max(np.abs(got - pub) / pub) < 1e-3
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b) * 1.01


def test_graph_replay_matches_eager_with_sparse_coarse():
    """CG(use_graph=True) captures blocks of iterations holding the sparse coarse sweeps;
    two graph solves (with eager work between them) equal the eager solve bit for bit."""
    print("DG S6 S4 S4 hybrid: graph replay vs eager, sparse coarse", flush=True)
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(6), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S4", "S4")
    out = []
    for graph in (False, True):
        s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(), dds.SparseCholesky(leaf_size=8, top_max=64)),
                   rtol=1e-9, maxiter=2000, use_graph=graph)
        s.setup(_torch_csr(A), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev),
                block_size=3)
        bt = torch.as_tensor(b, device=dev)
        for rep in range(2 if graph else 1):
            x, md = s.solve(bt)
            torch.cuda.synchronize()
            out.append((x.cpu().numpy(), md["iterations"]))
            _ = (bt * 2).sum().item()  # eager work between replays
    (xe, ie), (xg1, ig1), (xg2, ig2) = out
    assert ie == ig1 == ig2
    assert np.array_equal(xe, xg1) and np.array_equal(xe, xg2)
