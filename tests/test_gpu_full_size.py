"""The benchmark workloads at their full sizes, through size-independent properties (the
oracle cannot invert 4096 subdomains of 1000 unknowns in a test's time):

* Q1 64^3 (n = 63^3, which none of 2, 4, 8 divides): the FP64 solve with the BSR SpMV at
  b = 2, 4, 8 (padded operator, zero vector tail) against the same solve with the SELL SpMV;
* Q1 128^3 (n = 127^3, 4096 subdomains): the FP64 solve, and the FP16 / BF16 / FP32
  storage forms with FP32 accumulation against it;
* P1 512^2 jittered, 256 uneven k-means subdomains, 16 right-hand sides on the matrix cores.

Every solve is checked by its FP64 true residual ||b - A x|| / ||b|| (host, scipy) against
rtol 1e-8, its iteration count, and its distance to the FP64 solution.
"""

import numpy as np
import pytest
import torch

import dd_solvers as dds
from dd_solvers import problems as P

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")
RTOL = 1e-8


def _csr(A):
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                   torch.as_tensor(A.data), A.shape).to(dev)


def _solve(pb, A, storage=torch.float64, accumulate=None, bsr=None, layout="auto", b=None):
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(storage, scaled=storage == torch.float16,
                                                        accumulate=accumulate, layout=layout),
                                 dds.DenseCholesky()),
               rtol=RTOL, maxiter=2000, bsr_matmul=bsr is not None, check_every=8, use_graph=True)
    s.setup(A, decomposition=pb.decomposition, **({"block_size": bsr} if bsr else {}))
    rhs = pb.b if b is None else b
    x, m = s.solve(torch.as_tensor(rhs, device=dev))
    x = x.cpu().numpy()
    res = np.linalg.norm(rhs - pb.A @ x, axis=0) / np.linalg.norm(rhs, axis=0)
    layout_used = s.preconditioner.local.layout
    s.destroy()
    torch.cuda.empty_cache()
    return x, m, res, layout_used


@pytest.fixture(scope="module")
def q1_64():
    pb = P.make_problem("q1_3d", 64, box=8, overlap=1, seed=0, lognormal=True)
    A = _csr(pb.A)
    x, m, res, _ = _solve(pb, A)
    return pb, A, x, m


def test_q1_64_fp64(q1_64):
    pb, A, x, m = q1_64
    assert pb.A.shape[0] == 63 ** 3
    assert np.linalg.norm(pb.b - pb.A @ x) / np.linalg.norm(pb.b) < RTOL
    assert 40 <= m["iterations"] <= 55


@pytest.mark.parametrize("b", [2, 4, 8])
def test_q1_64_bsr_block_sizes_padded(q1_64, b):
    pb, A, x_sell, m_sell = q1_64
    n = pb.A.shape[0]
    assert n % b != 0
    print(f"64^3 BSR b={b}", flush=True)
    x, m, res, _ = _solve(pb, A, bsr=b)
    assert res < RTOL
    assert abs(m["iterations"] - m_sell["iterations"]) <= 1
    assert np.linalg.norm(x - x_sell) <= 1e-10 * np.linalg.norm(x_sell)


@pytest.fixture(scope="module")
def q1_128():
    pb = P.make_problem("q1_3d", 128, box=8, overlap=1, seed=0, lognormal=True)
    A = _csr(pb.A)
    x, m, res, layout = _solve(pb, A)
    assert layout == "sym"
    return pb, A, x, m, res


def test_q1_128_fp64(q1_128):
    pb, A, x, m, res = q1_128
    assert pb.A.shape[0] == 127 ** 3 and pb.decomposition.n_subdomains == 4096
    assert res < RTOL
    assert 45 <= m["iterations"] <= 55


@pytest.mark.parametrize("storage,accumulate,layout", [
    (torch.float16, torch.float32, "sym_mfma"), (torch.bfloat16, torch.float32, "sym_mfma"),
    (torch.float32, torch.float32, "sym"), (torch.float16, None, "sym")])
def test_q1_128_low_precision(q1_128, storage, accumulate, layout):
    """Low-precision local inverses (FP64 Krylov, coarse and residual check): the FP64
    residual meets rtol, at most one more iteration than FP64, and the solution within 1e-7
    of the FP64 one."""
    pb, A, x64, m64, _ = q1_128
    print(f"128^3 {storage} accumulate {accumulate}", flush=True)
    x, m, res, used = _solve(pb, A, storage=storage, accumulate=accumulate)
    assert used == layout
    assert res < RTOL
    assert abs(m["iterations"] - m64["iterations"]) <= 1
    assert np.linalg.norm(x - x64) <= 1e-7 * np.linalg.norm(x64)


@pytest.mark.parametrize("storage", [torch.float16, torch.bfloat16])
def test_p1_512_kmeans_16_rhs(storage):
    pb = P.make_problem("p1_2d_kmeans", 512, overlap=2, seed=0, nrhs=16, n_clusters=256, jitter=0.2)
    sizes = pb.decomposition.sizes
    assert sizes.shape[0] == 256 and sizes.min() < sizes.max()
    A = _csr(pb.A)
    print(f"512^2 k-means 16 RHS {storage}", flush=True)
    b0 = np.ascontiguousarray(pb.b[:, 0])
    x0, m0, res0, _ = _solve(pb, A, b=b0)
    assert res0 < RTOL
    X, m, res, used = _solve(pb, A, storage=storage, layout="mfma", b=np.ascontiguousarray(pb.b))
    assert used == "mfma"
    assert res.shape == (16,) and np.all(res < RTOL), res
    its = np.asarray(m["iterations"])
    assert np.all(its >= m0["iterations"] - 2) and np.all(its <= m0["iterations"] + 5), its
    assert np.linalg.norm(X[:, 0] - x0) <= 1e-7 * np.linalg.norm(x0)
