"""Pins of the CPU oracle: reference fixtures and known answers, closed forms, brute force.

Everything here runs on the CPU; the GPU path is compared with this oracle elsewhere.
"""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from helpers import bsr8, dg24_permuted
from oracle import kernels as K
from oracle.pcg import ConvergenceNotAchieved, cg
from oracle.schwarz import SchwarzOracle

from dd_solvers import problems as P


# ---------------------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("k", [1, 2, 5, 10, 20, 50])
def test_gemv_matches_matmul(k):
    # shapes of src/dd_solvers/tests/test_gemv.py:22-35
    rng = np.random.default_rng(42)
    A = rng.standard_normal((7, k, k))
    b = rng.standard_normal((7, k))
    out = K.gemv_strided_batched(A, b)
    assert np.allclose(out, np.einsum("nrc,nc->nr", A, b), rtol=1e-13, atol=1e-13)


def test_gemv_accumulates_in_vector_dtype():
    A = np.ones((1, 3, 3), dtype=np.float32)
    b = np.array([[1e8, 1.0, -1e8]], dtype=np.float32)
    # ascending-c accumulation in float32 loses the 1.0: (1e8 + 1) - 1e8 = 0
    assert K.gemv_strided_batched(A, b)[0, 0] == 0.0
    assert K.gemv_strided_batched(A, b.astype(np.float64))[0, 0] == 1.0


def test_gemv_shape_error():
    with pytest.raises(RuntimeError):
        K.gemv_strided_batched(np.zeros((2, 3, 3)), np.zeros((2, 4)))


def test_bsr_fixture_matches_dense():
    crow, col, vals = bsr8()  # src/dd_solvers/tests/test_bsr.py:9-138
    b = vals.shape[1]
    nbr = crow.shape[0] - 1
    dense = np.zeros((nbr * b, nbr * b))
    for br in range(nbr):
        for blk in range(crow[br], crow[br + 1]):
            dense[br * b:(br + 1) * b, col[blk] * b:(col[blk] + 1) * b] += vals[blk]
    x = np.random.default_rng(0).standard_normal(nbr * b)
    assert np.allclose(K.bsrmv(crow, col, vals, x), dense @ x, atol=1e-12)


def test_csr_to_bsr_known_answer():
    # src/dd_solvers/tests/test_bsr.py:196-223
    indptr = [0, 2, 3, 5, 9]
    indices = [0, 1, 1, 0, 2, 0, 1, 2, 3]
    data = np.array([1, 2, 3, 4, 5, 6, 7, 8, 9], dtype=np.float64)
    crow, col, vals = K.csr_to_bsr(indptr, indices, data, 2, (4, 4))
    assert crow.tolist() == [0, 1, 3]
    assert col.tolist() == [0, 0, 1]
    assert np.array_equal(vals, [[[1, 2], [0, 3]], [[4, 0], [6, 7]], [[5, 0], [8, 9]]])


def test_csr_to_bsr_round_trip_fixture():
    crow, col, vals = bsr8()
    b = vals.shape[1]
    nbr = crow.shape[0] - 1
    dense = np.zeros((nbr * b, nbr * b))
    for br in range(nbr):
        for blk in range(crow[br], crow[br + 1]):
            dense[br * b:(br + 1) * b, col[blk] * b:(col[blk] + 1) * b] += vals[blk]
    A = sps.csr_matrix(dense)
    c2, col2, v2 = K.csr_to_bsr(A.indptr, A.indices, A.data, b, A.shape)
    back = np.zeros_like(dense)
    for br in range(nbr):
        for blk in range(c2[br], c2[br + 1]):
            back[br * b:(br + 1) * b, col2[blk] * b:(col2[blk] + 1) * b] += v2[blk]
    assert np.array_equal(back, dense)


def test_sort_csr_known_answer():
    # src/dd_solvers/tests/test_cusparse.py:8-34
    ip, ix, v = K.sort_csr([0, 2, 3, 5, 9], [1, 0, 1, 0, 2, 2, 3, 1, 0],
                           np.array([2, 1, 3, 4, 5, 8, 9, 7, 6], dtype=np.float64))
    assert ip.tolist() == [0, 2, 3, 5, 9]
    assert ix.tolist() == [0, 1, 1, 0, 2, 0, 1, 2, 3]
    assert v.tolist() == [1, 2, 3, 4, 5, 6, 7, 8, 9]
    assert ix.dtype == np.int32


@pytest.mark.parametrize("case", [
    # src/dd_solvers/tests/test_cusparse.py:37-78
    (([0, 3, 6], [0, 2, 2, 0, 0, 1], [1.0, 2, 3, 4, 5, 6]), ([0, 2, 4], [0, 2, 0, 1], [1.0, 5, 9, 6])),
    (([0, 2, 2, 5], [0, 1, 0, 0, 1], [1.0, 2, 3, 4, 5]), ([0, 2, 2, 4], [0, 1, 0, 1], [1.0, 2, 7, 5])),
])
def test_gather_csr_known_answers(case):
    (ip, ix, v), (eip, eix, ev) = case
    oip, oix, ov = K.gather_csr(ip, ix, np.array(v))
    assert oip.tolist() == eip and oix.tolist() == eix and ov.tolist() == ev


def test_spd_solve_known_answers():
    # src/dd_solvers/tests/test_cudss.py:15-84
    A = np.array([[2.0, 0, 2], [0, 2, 0], [2, 0, 3]])
    assert np.allclose(K.spd_solve(A, [2, -2, 2]), [1, -1, 0], atol=1e-12)
    assert np.allclose(K.spd_solve(A, [4, 4, 5]), [1, 2, 1], atol=1e-12)
    B = np.array([[2.0, 1], [1, 1]])
    x = np.concatenate([K.spd_solve(A, [4, -2, 5]), K.spd_solve(B, [-1, 0])])
    assert np.allclose(x, [1, -1, 1, -1, 1], atol=1e-12)


def test_round_bf16():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 65504.0], dtype=np.float32)
    r = K.round_bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == np.float32(-3.140625)
    assert np.all(K.round_bf16(r) == r)  # idempotent


# ---------------------------------------------------------------------------------------
# CG
# ---------------------------------------------------------------------------------------

def test_cg_condition_estimate_known_answer():
    # src/dd_solvers/tests/test_solvers.py:179-184
    A = np.diag([0.5, 4.0, 21.0])
    _, meta = cg(lambda v: A @ v, np.ones(3), rtol=1e-9, estimate_cond=True)
    assert np.isclose(meta["condition number estimate"], 42.0)


def test_cg_unpreconditioned_matches_direct_and_terminates():
    A = P.assemble_q1_3d(6)  # n = 125
    b = P.uniform_rhs(A.shape[0], 3)
    x, meta = cg(lambda v: A @ v, b, rtol=1e-12, maxiter=None, record_iterates=True)
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-10
    assert meta["iterations"] <= A.shape[0]
    # the recurrence residual follows the true residual b - A x_i
    it = meta["iterates"]
    norms = meta["residual norms"]
    for i in (0, len(it) // 2, len(it) - 1):
        assert abs(np.linalg.norm(b - A @ it[i]) - norms[i]) <= 1e-8 * np.linalg.norm(b)


def test_cg_exact_preconditioner_one_iteration():
    A = P.assemble_q1_3d(5)
    Ad = A.toarray()
    b = P.uniform_rhs(A.shape[0], 1)
    Ainv = np.linalg.inv(Ad)
    _, meta = cg(lambda v: A @ v, b, precond=lambda r: Ainv @ r, rtol=1e-10)
    assert meta["iterations"] == 1


def test_cg_raise_quirk():
    A = np.diag([1.0, 2.0, 3.0, 4.0])
    b = np.ones(4)
    with pytest.raises(ConvergenceNotAchieved) as e:
        cg(lambda v: A @ v, b, rtol=1e-12, maxiter=2)
    assert e.value.metadata["iterations"] == 1
    # converges at check 4 (i = 4); maxiter = 5 -> i + 1 == maxiter raises anyway
    _, meta = cg(lambda v: A @ v, b, rtol=1e-12, maxiter=6)
    assert meta["iterations"] == 4
    with pytest.raises(ConvergenceNotAchieved):
        cg(lambda v: A @ v, b, rtol=1e-12, maxiter=5)


# ---------------------------------------------------------------------------------------
# Schwarz
# ---------------------------------------------------------------------------------------

def _uniform(Ap, k, spc, **kw):
    n = Ap.shape[0]
    ptr = np.arange(0, n + 1, k)
    return SchwarzOracle(Ap, ptr, ptr, np.arange(n), spc, **kw)


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
@pytest.mark.parametrize("storage", ["float64", "float32", "float16", "bfloat16"])
def test_dg24_fixture_solves(kind, storage):
    # src/dd_solvers/tests/test_solvers.py:149-176: ||A x - b|| < 1e-10 after unpermuting
    A, b, perm, Ap, k, spc = dg24_permuted()
    M = _uniform(Ap, k, spc, kind=kind, storage=storage)
    T0 = M.T0 if kind == "hybrid" else None
    x, meta = cg(lambda v: Ap @ v, b[perm], precond=M.apply, T0=T0, rtol=1e-7)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.shape[0])
    assert np.linalg.norm(A @ x[inv] - b) < 1e-10


def _small_problem(overlap=1):
    return P.make_problem("q1_3d", 6, box=2, overlap=overlap, seed=5, lognormal=True)


@pytest.mark.parametrize("overlap", [0, 1])
def test_additive_is_spd(overlap):
    pb = _small_problem(overlap)
    d = pb.decomposition
    M = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, kind="additive").dense()
    assert np.allclose(M, M.T, atol=1e-12)
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_hybrid_projection_identities():
    pb = _small_problem(1)
    d = pb.decomposition
    S = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, kind="hybrid")
    A = pb.A.toarray()
    RC = S.RC.toarray()
    P_ = np.eye(A.shape[0]) - RC.T @ np.linalg.solve(RC @ A @ RC.T, RC @ A)
    assert np.allclose(P_ @ P_, P_, atol=1e-10)
    assert np.allclose(A @ P_, P_.T @ A, atol=1e-10)
    # the hybrid apply equals P (sum R_i^T A_i^-1 R_i) on any vector
    r = np.random.default_rng(1).standard_normal(A.shape[0])
    assert np.allclose(S.apply(r), P_ @ S.local(r), atol=1e-10)
    # T0 b is the coarse Galerkin projection of b
    b = pb.b
    assert np.allclose(S.T0(b), RC.T @ np.linalg.solve(RC @ A @ RC.T, RC @ b), atol=1e-10)


def test_hybrid_planted_sign_error_is_caught():
    """A flipped coarse-correction sign breaks the projection identity above."""
    pb = _small_problem(1)
    d = pb.decomposition
    S = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, kind="hybrid")
    A = pb.A.toarray()
    RC = S.RC.toarray()
    r = np.random.default_rng(1).standard_normal(A.shape[0])
    y = S.local(r)
    wrong = y + S.prolong(S.coarse_solve(S.R0A @ y))
    P_ = np.eye(A.shape[0]) - RC.T @ np.linalg.solve(RC @ A @ RC.T, RC @ A)
    assert not np.allclose(wrong, P_ @ y, atol=1e-6)


def test_single_subdomain_is_exact():
    pb = _small_problem(0)
    n = pb.A.shape[0]
    S = SchwarzOracle(pb.A, [0, n], [0, n], np.arange(n), kind="additive", coarse=False)
    _, meta = cg(lambda v: pb.A @ v, pb.b, precond=S.apply, rtol=1e-10)
    assert meta["iterations"] == 1


def test_point_subdomains_without_coarse_is_jacobi():
    pb = _small_problem(0)
    n = pb.A.shape[0]
    ptr = np.arange(n + 1)
    S = SchwarzOracle(pb.A, ptr, ptr, np.arange(n), kind="additive", coarse=False)
    r = np.random.default_rng(2).standard_normal(n)
    assert np.allclose(S.apply(r), r / pb.A.diagonal(), rtol=1e-13)


def test_overlap_apply_brute_force():
    """sum_i R_i^T A_i^-1 R_i r with explicit 0/1 restriction matrices."""
    pb = _small_problem(1)
    d = pb.decomposition
    S = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, kind="additive", coarse=False)
    A = pb.A.toarray()
    n = A.shape[0]
    r = np.random.default_rng(3).standard_normal(n)
    y = np.zeros(n)
    for i in range(d.n_subdomains):
        idx = d.subdomain(i)
        R = np.zeros((idx.shape[0], n))
        R[np.arange(idx.shape[0]), idx] = 1.0
        y += R.T @ np.linalg.solve(R @ A @ R.T, R @ r)
    assert np.allclose(S.apply(r), y, atol=1e-12)


def test_scaled_storage():
    """SchwarzOracle(scaled=True) stores A_i^-1 / s_i with s_i = float32(max |A_i^-1|) and
    multiplies the product by s_i. FP64 storage: the same apply as unscaled to rounding.
    FP16 storage of an operator whose inverses exceed the FP16 range (A scaled by 1e-6, so
    |A_i^-1| ~ 1e6 > 65504): unscaled storage overflows, scaled storage keeps FP16's
    relative precision. A scale that is stored but not applied back is caught."""
    pb = _small_problem(1)
    d = pb.decomposition
    r = np.random.default_rng(4).standard_normal(pb.A.shape[0])
    plain = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False)
    sc = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False, scaled=True)
    assert np.allclose(sc.local(r), plain.local(r), rtol=1e-14, atol=1e-14 * np.abs(plain.local(r)).max())
    for i, (idx, Aq) in enumerate(sc.inv):
        assert np.abs(Aq).max() == pytest.approx(1.0, rel=1e-7)  # max |A^-1 / s| is 1
        inv_i = np.linalg.inv(pb.A[idx][:, idx].toarray())
        assert sc.scale[i] == float(np.float32(np.abs(inv_i).max()))
    small = pb.A * 1e-6
    exact = SchwarzOracle(small, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False).local(r)
    with np.errstate(over="ignore", invalid="ignore"):
        over = SchwarzOracle(small, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False,
                             storage="float16").local(r)
    assert not np.all(np.isfinite(over))
    h = SchwarzOracle(small, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False, storage="float16",
                      scaled=True)
    y = h.local(r)
    assert np.linalg.norm(y - exact) <= 2e-3 * np.linalg.norm(exact)
    h.scale = [1.0] * len(h.scale)  # planted: the scale not applied back
    assert np.linalg.norm(h.local(r) - exact) > 0.5 * np.linalg.norm(exact)


def test_given_inverses():
    """SchwarzOracle(inverses=...) applies the given matrices as they are: with the
    oracle's own FP64 inverses it reproduces the default apply; with any other matrices
    it computes sum_i R_i^T M_i R_i r by brute force."""
    pb = _small_problem(1)
    d = pb.decomposition
    r = np.random.default_rng(5).standard_normal(pb.A.shape[0])
    ref = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False)
    given = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False,
                          inverses=[m for _, m in ref.inv])
    assert np.array_equal(given.local(r), ref.local(r))
    rng = np.random.default_rng(6)
    Ms = [rng.standard_normal((int(k), int(k))) for k in np.diff(d.sub_ptr)]
    g2 = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, coarse=False, inverses=Ms)
    y = np.zeros_like(r)
    for i, M in enumerate(Ms):
        idx = d.subdomain(i)
        y[idx] += M @ r[idx]
    assert np.allclose(g2.local(r), y, rtol=1e-13, atol=1e-13)


def test_cho_solve_matches_dense_inverse():
    rng = np.random.default_rng(4)
    B = rng.standard_normal((40, 40))
    A = B @ B.T + 40 * np.eye(40)
    x = rng.standard_normal(40)
    assert np.allclose(K.spd_solve(A, x), np.linalg.inv(A) @ x, rtol=1e-12)
    L = K.cholesky_lower(A)
    assert np.allclose(L @ L.T, A, atol=1e-10)
    assert np.allclose(sla.solve_triangular(L.T, sla.solve_triangular(L, x, lower=True)),
                       K.spd_solve(A, x), rtol=1e-12)


# ---------------------------------------------------------------------------------------
# adjoint gradient
# ---------------------------------------------------------------------------------------

def _c5_inputs(N, seed=0):
    theta = 0.5 * np.random.default_rng([seed, 1]).standard_normal(N ** 3)
    theta_obs = 0.5 * np.random.default_rng([seed + 1, 1]).standard_normal(N ** 3)
    f = P.uniform_rhs((N - 1) ** 3, seed)
    return theta, theta_obs, f


def test_element_gradient_matches_brute_force():
    from oracle.adjoint import element_gradient
    N = 4
    theta, _, f = _c5_inputs(N)
    kappa = np.exp(theta).reshape(N, N, N)
    rng = np.random.default_rng(5)
    lam = rng.standard_normal((N - 1) ** 3)
    u = rng.standard_normal((N - 1) ** 3)
    g = element_gradient(N, kappa, lam, u)
    for e in (0, 7, 21, N ** 3 - 1):
        d = np.zeros(N ** 3)
        d[e] = kappa.reshape(-1)[e]  # dA/dtheta_e = k_e * dA/dk_e, A linear in k
        dA = P.assemble_q1_3d(N, d.reshape(N, N, N)).toarray()
        assert np.isclose(g[e], -lam @ dA @ u, rtol=1e-12, atol=1e-14)


def test_adjoint_gradient_matches_central_differences():
    from oracle.adjoint import fd_directional, objective_and_gradient
    N = 8
    theta, theta_obs, f = _c5_inputs(N)
    import scipy.sparse.linalg as spla
    u_obs = spla.spsolve(P.assemble_q1_3d(N, np.exp(theta_obs).reshape(N, N, N)).tocsc(), f)
    J, g, u, lam = objective_and_gradient(N, theta, f, u_obs)
    dirs = np.random.default_rng(7).standard_normal((32, N ** 3))
    fd = fd_directional(N, theta, f, u_obs, dirs, eps=1e-6)
    ad = dirs @ g
    assert np.all(np.abs(fd - ad) <= 1e-6 * np.abs(ad).max())


def _blocked_dot(u, v):
    """u . v summed in another order: 256 strided partial sums, added in reverse."""
    w = np.asarray(u) * np.asarray(v)
    m = -(-w.shape[0] // 256) * 256
    return float(np.pad(w, (0, m - w.shape[0])).reshape(-1, 256).sum(axis=0)[::-1].sum())


def _reversed_rows(A):
    """The same CSR matrix with each row's entries in reverse order (the SpMV sums them in
    reverse)."""
    A = A.tocsr()
    data, idx = A.data.copy(), A.indices.copy()
    for i in range(A.shape[0]):
        a, b = A.indptr[i], A.indptr[i + 1]
        data[a:b], idx[a:b] = data[a:b][::-1], idx[a:b][::-1]
    B = sps.csr_matrix((data, idx, A.indptr.copy()), shape=A.shape)
    B.has_sorted_indices = False
    return B


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_dg_cg_reordered_reduction_drift(kind):
    """How far two FP64 runs of the same PCG drift apart when only their summation orders
    differ (dots, norms, SpMV rows), on the DG p = 3 S5 / S3 / S2 system of the GPU DG
    tests: the measured envelope that their GPU-vs-oracle tolerances rest on. The first
    iterates agree to rounding; the residual histories separate after ~60-100 iterations;
    the iteration counts may differ by ~1% (259 vs 256 for Additive); the solutions agree
    far below the solver tolerance."""
    from dd_solvers import dg
    from oracle.schwarz import BlockSchwarzOracle
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(5), 3)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S3", "S2")
    O = BlockSchwarzOracle(A, k, spc, kind=kind)
    T0 = O.T0 if kind == "hybrid" else None
    x1, m1 = cg(lambda v: A @ v, b, precond=O.apply, T0=T0, rtol=1e-9, maxiter=2000)
    Ar = _reversed_rows(A)
    x2, m2 = cg(lambda v: Ar @ v, b, precond=O.apply, T0=T0, rtol=1e-9, maxiter=2000,
                dot=_blocked_dot)
    r1, r2 = np.array(m1["residual norms"]), np.array(m2["residual norms"])
    n = min(len(r1), len(r2))
    rel = np.abs(r1[:n] - r2[:n]) / r1[:n]
    assert rel[:10].max() < 1e-13
    assert rel[:50].max() < 1e-10
    assert abs(m1["iterations"] - m2["iterations"]) <= 0.02 * m1["iterations"]
    assert np.linalg.norm(x1 - x2) <= 1e-9 * np.linalg.norm(x1)
    if kind == "additive":  # the drift is real: the histories separate and the counts differ
        assert rel.max() > 1e-3 and m1["iterations"] != m2["iterations"]
