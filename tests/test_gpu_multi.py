"""GPU parity of the batched form: blocks of 16 right-hand sides, uneven subdomains, the
local inverses applied as an FP16/BF16 batched GEMM on the matrix cores.

Bars: one preconditioner application within 1e-5 (norm-wise, per column) of the oracle's
FP64 product with the same quantised, scaled inverses (the block path splits each column
into 2 FP16 or 3 BF16 terms and accumulates in FP32); per-column CG iteration counts within
+-2 of the oracle's CG with those inverses, solutions within 1e-6, and every column's true
FP64 residual below rtol; a column solved alone equals the same column of a block bit for
bit (columns never mix); graph replay equals eager execution bit for bit; every autotuner
launch shape gives the same bits.
"""

import functools

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from helpers import dg24_permuted
from oracle.pcg import cg as oracle_cg
from oracle.schwarz import SchwarzOracle

import dd_solvers as dds
from dd_solvers import problems as P
from dd_solvers.tuning import autotune_mfma, mfma_candidates

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")
TD = {"float16": torch.float16, "bfloat16": torch.bfloat16}


def _torch_csr(A):
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                   torch.as_tensor(A.data), A.shape).to(device=dev, dtype=torch.float64)


@functools.lru_cache(maxsize=None)
def _kmeans(nrhs=16):
    # jittered P1 on 96^2, 12 k-means subdomains (uneven, ~600-1000 unknowns), overlap 2
    return P.make_problem("p1_2d_kmeans", 96, overlap=2, seed=5, nrhs=nrhs, n_clusters=12, jitter=0.2)


@functools.lru_cache(maxsize=None)
def _q1_nonoverlap():
    return P.make_problem("q1_3d", 17, box=4, overlap=0, seed=2, lognormal=True, nrhs=16)


def _oracle(pb, kind, storage):
    d = pb.decomposition
    return SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse, kind=kind,
                         storage=storage, scaled=True)


def _pc(kind, storage, **kw):
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    return cls(torch.float64, dds.Inv(TD[storage], scaled=True, layout="mfma", **kw), dds.DenseCholesky())


def _colrel(a, b):
    return np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_block_apply_matches_oracle(kind, storage):
    pb = _kmeans()
    pc = _pc(kind, storage)
    pc.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    R = np.random.default_rng(11).uniform(-1, 1, (pb.A.shape[0], 16))
    R[:, 3] *= 1e-6  # a column far below the others: per-column scaling keeps its accuracy
    Z, _ = pc.solve(torch.as_tensor(R, device=dev))
    O = _oracle(pb, kind, storage)
    Zo = np.stack([O.apply(R[:, j]) for j in range(16)], axis=1)
    rel = _colrel(Z.cpu().numpy(), Zo)
    assert rel.max() <= 1e-5, rel


def test_block_apply_small_subdomains_dg24():
    """Subdomains of 3 unknowns: one 16-row block, one 32-column pair, mostly padding."""
    A, b, perm, Ap, k, spc = dg24_permuted()
    n = Ap.shape[0]
    ptr = np.arange(0, n + 1, k)
    pc = _pc("hybrid", "float16")
    pc.setup(_torch_csr(Ap), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev))
    R = np.random.default_rng(1).uniform(-1, 1, (n, 16))
    Z, _ = pc.solve(torch.as_tensor(R, device=dev))
    O = SchwarzOracle(Ap, ptr, ptr, np.arange(n), spc, kind="hybrid", storage="float16", scaled=True)
    Zo = np.stack([O.apply(R[:, j]) for j in range(16)], axis=1)
    assert _colrel(Z.cpu().numpy(), Zo).max() <= 1e-5


def _block_parity(pb, kind, storage, m, rtol=1e-8, **cg_kw):
    O = _oracle(pb, kind, storage)
    s = dds.CG(_pc(kind, storage), rtol=rtol, bsr_matmul=False, **cg_kw)
    s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    B = np.ascontiguousarray(pb.b[:, :m])
    X, meta = s.solve(torch.as_tensor(B, device=dev))
    X = X.cpu().numpy()
    assert X.shape == (pb.A.shape[0], m)
    assert len(meta["iterations"]) == m
    for j in range(m):
        xo, mo = oracle_cg(lambda v: pb.A @ v, B[:, j], precond=O.apply,
                           T0=O.T0 if kind == "hybrid" else None, rtol=rtol)
        assert abs(meta["iterations"][j] - mo["iterations"]) <= 2, (j, meta["iterations"][j], mo["iterations"])
        err = np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo)
        assert err <= 1e-6, (j, err)
        res = np.linalg.norm(B[:, j] - pb.A @ X[:, j]) / np.linalg.norm(B[:, j])
        assert res < rtol * 1.01, (j, res)
        assert len(meta["residual norms"][j]) == meta["iterations"][j] + 1
    return X, meta, s


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
def test_block_cg_hybrid_kmeans(storage):
    # 13 of 16 columns: three padding columns must stay inert
    _block_parity(_kmeans(), "hybrid", storage, 13)


def test_block_cg_additive_two_blocks():
    # 20 columns: a full block of 16, then a block of 4
    pb = _kmeans(nrhs=20)
    _block_parity(pb, "additive", "float16", 20)


def test_block_cg_nonoverlapping_boxes():
    _block_parity(_q1_nonoverlap(), "hybrid", "bfloat16", 16)


def test_single_column_equals_block_column():
    pb = _kmeans()
    s = dds.CG(_pc("hybrid", "float16"), rtol=1e-8, bsr_matmul=False)
    s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    B = torch.as_tensor(np.ascontiguousarray(pb.b), device=dev)
    X, meta = s.solve(B)
    x1, m1 = s.solve(B[:, 5].contiguous())
    assert x1.dim() == 1
    assert m1["iterations"] == meta["iterations"][5]
    assert torch.equal(x1, X[:, 5])


def test_block_graph_replay_matches_eager():
    pb = _kmeans()
    B = torch.as_tensor(np.ascontiguousarray(pb.b), device=dev)
    out = []
    for use_graph in (False, True):
        s = dds.CG(_pc("hybrid", "float16"), rtol=1e-8, bsr_matmul=False, check_every=4,
                   use_graph=use_graph)
        s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
        X1, m1 = s.solve(B)
        X2, m2 = s.solve(B)  # second replay of the same graph
        assert torch.equal(X1, X2) and m1["iterations"] == m2["iterations"]
        out.append((X1, m1))
    assert out[0][1]["iterations"] == out[1][1]["iterations"]
    assert torch.equal(out[0][0], out[1][0])


def test_autotuner_shapes_are_bitwise_equal():
    pb = _kmeans()
    pc = _pc("additive", "float16")
    pc.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    X = torch.as_tensor(np.random.default_rng(3).uniform(-1, 1, (pb.A.shape[0], 16)), device=dev)
    mf = pc._mfma
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    ref = None
    shapes = mfma_candidates(pc.local.n_items, n_cu)
    assert any(g < 0 for g, _, _ in shapes) and any(g >= 0 for g, _, _ in shapes)
    for shape in shapes:
        mf["grid"], mf["waves"], mf["rpw"] = shape
        Y = pc._local_multi_into(X, None).clone()
        if ref is None:
            ref = Y
        assert torch.equal(Y, ref), shape
    best = autotune_mfma(pc)
    assert (best["grid"], best["waves"], best["rpw"]) in best["timings"]
    assert (mf["grid"], mf["waves"], mf["rpw"]) == (best["grid"], best["waves"], best["rpw"])
    assert torch.equal(pc._local_multi_into(X, None), ref)


def test_autotune_flag_in_setup():
    pb = _kmeans()
    s = dds.CG(_pc("hybrid", "bfloat16", autotune=True), rtol=1e-8, bsr_matmul=False)
    s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    assert s.preconditioner.tuning["waves"] in (1, 2, 4, 8)
    X, meta = s.solve(torch.as_tensor(np.ascontiguousarray(pb.b[:, :4]), device=dev))
    A = pb.A
    for j in range(4):
        assert np.linalg.norm(pb.b[:, j] - A @ X[:, j].cpu().numpy()) < 1.01e-8 * np.linalg.norm(pb.b[:, j])


def test_block_requires_mfma_layout():
    pb = _kmeans()
    s = dds.CG(dds.HybridSchwarz(torch.float64, dds.Inv(torch.float16), dds.DenseCholesky()),
               bsr_matmul=False)
    s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    with pytest.raises(ValueError):
        s.solve(torch.as_tensor(np.ascontiguousarray(pb.b[:, :2]), device=dev))


def test_block_sellmm_and_dots_match_host():
    """SpMM through the SELL kernel, with the fused per-column p . q partials."""
    pb = _kmeans()
    s = dds.CG(None, bsr_matmul=False)
    s.setup(_torch_csr(pb.A))
    n = pb.A.shape[0]
    X = np.random.default_rng(4).uniform(-1, 1, (n, 16))
    Xt = torch.as_tensor(X, device=dev)
    Y = torch.empty_like(Xt)
    s.matrix._enqueue_matmat(Xt, Y, None, None)
    ref = sps.csr_matrix(pb.A) @ X
    assert np.allclose(Y.cpu().numpy(), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_block_x0_mutated_and_convergence_not_achieved():
    # Additive: the reference's Hybrid start x = x0 + T0(b) (solvers.py:233-234) leaves a
    # residual outside the space the balancing preconditioner is SPD on unless x0 = 0
    pb = _kmeans()
    s = dds.CG(_pc("additive", "float16"), rtol=1e-8, bsr_matmul=False)
    s.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    B = torch.as_tensor(np.ascontiguousarray(pb.b[:, :3]), device=dev)
    X, m = s.solve(B)
    # a nonzero initial guess is used and, as in the 1-D path, mutated and returned
    x0 = 0.5 * X.clone()
    out, m0 = s.solve(B, x0)
    assert out is x0
    assert max(m0["iterations"]) <= max(m["iterations"])
    for j in range(3):
        r = np.linalg.norm(pb.b[:, j] - pb.A @ x0[:, j].cpu().numpy()) / np.linalg.norm(pb.b[:, j])
        assert r < 1.01e-8
    s2 = dds.CG(_pc("hybrid", "float16"), rtol=1e-12, maxiter=3, bsr_matmul=False)
    s2.setup(_torch_csr(pb.A), decomposition=pb.decomposition)
    with pytest.raises(dds.ConvergenceNotAchieved) as e:
        s2.solve(B)
    assert all(len(r) == 3 for r in e.value.metadata["residual norms"])


def test_block_condition_number_estimate():
    """Per-column Lanczos estimate from the device-side alpha/beta logs: a diagonal SPD
    matrix with known eigenvalues, unpreconditioned (the reference's known answer,
    test_solvers.py:179-184, for every column)."""
    A = sps.diags([0.5, 4.0, 21.0]).tocsr()
    s = dds.CG(preconditioner=None, rtol=1e-9, estimate_cond=True, bsr_matmul=False)
    s.setup(_torch_csr(A))
    B = torch.ones((3, 5), dtype=torch.float64, device=dev)
    _, meta = s.solve(B)
    assert len(meta["condition number estimate"]) == 5
    assert np.allclose(meta["condition number estimate"], 42.0)


# -- 16 right-hand sides on the symmetric matrix-core tiles (dds_sym_mfma_apply16) ---------
@functools.lru_cache(maxsize=None)
def _kmeans13():
    # the same grid in 13 k-means subdomains: 720-951 unknowns (k <= 1024 for this path)
    return P.make_problem("p1_2d_kmeans", 96, overlap=2, seed=5, nrhs=16, n_clusters=13, jitter=0.2)


def _pc_sym(kind, storage):
    cls = dds.AdditiveSchwarz if kind == "additive" else dds.HybridSchwarz
    return cls(torch.float64, dds.Inv(TD[storage], scaled=True, layout="sym_mfma",
                                      accumulate=torch.float32), dds.DenseCholesky())


def _stored_oracle(pb_A, d, kind, pc, spc=None):
    return SchwarzOracle(pb_A, d.owner_ptr, d.sub_ptr, d.sub_idx,
                         d.solvers_per_coarse if spc is None else spc, kind=kind,
                         inverses=pc.stored_local_inverses())


@pytest.mark.parametrize("storage", ["float16", "bfloat16"])
@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_sym16_block_apply_matches_stored_inverses(kind, storage):
    """The symmetric store applied to 16 columns: within 1e-5 per column of the FP64 product
    of the very inverses it stores (16-bit terms of x / sigma, FP32 accumulation), the local
    part alone and the whole preconditioner."""
    pb = _kmeans13()
    d = pb.decomposition
    pc = _pc_sym(kind, storage)
    pc.setup(_torch_csr(pb.A), decomposition=d)
    assert pc.local.layout == "sym_mfma" and pc._msym is not None and pc._mfma is None
    print(f"sym16 apply: {kind} {storage} k {np.diff(d.sub_ptr).min()}..{np.diff(d.sub_ptr).max()}",
          flush=True)
    R = np.random.default_rng(11).uniform(-1, 1, (pb.A.shape[0], 16))
    R[:, 3] *= 1e-6  # a column far below the others: per-column scaling keeps its accuracy
    Rt = torch.as_tensor(R, device=dev)
    O = _stored_oracle(pb.A, d, kind, pc)
    yl = pc._local_multi_into(Rt.clone(), None).cpu().numpy()
    ylo = np.stack([O.local(R[:, j]) for j in range(16)], axis=1)
    assert _colrel(yl, ylo).max() <= 1e-5
    Z, _ = pc.solve(Rt)
    Zo = np.stack([O.apply(R[:, j]) for j in range(16)], axis=1)
    assert _colrel(Z.cpu().numpy(), Zo).max() <= 1e-5


@pytest.mark.parametrize("case", ["dg24", "q1_boxes"])
def test_sym16_small_and_partial_tiles(case):
    """Subdomains of 3 unknowns (one partial 64-row block) and of 64 and fewer (boxes of the
    non-overlapping Q1 grid): the padding rows and columns of the tiles stay inert."""
    if case == "dg24":
        A, b, perm, Ap, k, spc = dg24_permuted()
        n = Ap.shape[0]
        ptr = np.arange(0, n + 1, k)
        pc = _pc_sym("hybrid", "float16")
        pc.setup(_torch_csr(Ap), dofs_per_solver=k, solvers_per_coarse=torch.as_tensor(spc, device=dev))
        O = SchwarzOracle(Ap, ptr, ptr, np.arange(n), spc, kind="hybrid",
                          inverses=pc.stored_local_inverses())
    else:
        pb = _q1_nonoverlap()
        Ap, d = pb.A, pb.decomposition
        n = Ap.shape[0]
        pc = _pc_sym("hybrid", "bfloat16")
        pc.setup(_torch_csr(Ap), decomposition=d)
        O = _stored_oracle(Ap, d, "hybrid", pc)
    assert pc._msym is not None
    R = np.random.default_rng(1).uniform(-1, 1, (n, 16))
    Z, _ = pc.solve(torch.as_tensor(R, device=dev))
    Zo = np.stack([O.apply(R[:, j]) for j in range(16)], axis=1)
    assert _colrel(Z.cpu().numpy(), Zo).max() <= 1e-5


def test_sym16_block_cg_and_graph_replay():
    """CG over a block of 13 columns through the symmetric store: per column the oracle CG's
    iteration count (+-2) with the stored inverses, the solution within 1e-6, the true FP64
    residual below rtol; a graph replay gives the eager bits; a single vector keeps the
    single-vector path."""
    pb = _kmeans13()
    d = pb.decomposition
    B = np.ascontiguousarray(pb.b[:, :13])
    out = []
    for use_graph in (False, True):
        s = dds.CG(_pc_sym("hybrid", "float16"), rtol=1e-8, bsr_matmul=False, check_every=4,
                   use_graph=use_graph)
        s.setup(_torch_csr(pb.A), decomposition=d)
        X, meta = s.solve(torch.as_tensor(B, device=dev))
        out.append((X, meta))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1]["iterations"] == out[1][1]["iterations"]
    X, meta = out[0][0].cpu().numpy(), out[0][1]
    O = _stored_oracle(pb.A, d, "hybrid", s.preconditioner)
    for j in range(13):
        xo, mo = oracle_cg(lambda v: pb.A @ v, B[:, j], precond=O.apply, T0=O.T0, rtol=1e-8)
        assert abs(meta["iterations"][j] - mo["iterations"]) <= 2, (j, meta["iterations"][j], mo["iterations"])
        assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) <= 1e-6
        assert np.linalg.norm(B[:, j] - pb.A @ X[:, j]) < 1.01e-8 * np.linalg.norm(B[:, j])
    x1, m1 = s.solve(torch.as_tensor(B[:, 0].copy(), device=dev))
    assert x1.dim() == 1 and abs(m1["iterations"] - meta["iterations"][0]) <= 2


def test_sym16_uneven_sizes_up_to_1024():
    """Subdomains of 1, 7, 63, 64, 65, 129, 1000 and 1024 unknowns (one node of overlap on
    each side) of a 1-D SPD operator: every last-tile shape (rows and column pairs of the
    partial 64-row block) and the largest subdomain the path takes, against the oracle's
    FP64 product with the stored inverses."""
    sizes = [1024, 65, 1, 1000, 7, 64, 129, 63]
    n = int(sum(sizes))
    main = 2.0 + np.random.default_rng(5).uniform(0.1, 1.0, n)
    A = sps.diags([-np.ones(n - 1), main, -np.ones(n - 1)], [-1, 0, 1], format="csr")
    owner_ptr = np.concatenate([[0], np.cumsum(sizes)])
    subs = [np.arange(max(0, a - 1), min(n, b + 1)) for a, b in zip(owner_ptr[:-1], owner_ptr[1:])]
    sub_ptr = np.concatenate([[0], np.cumsum([len(x) for x in subs])])
    d = P.Decomposition(perm=np.arange(n), owner_ptr=owner_ptr, sub_ptr=sub_ptr,
                        sub_idx=np.concatenate(subs), solvers_per_coarse=np.array([4, 4]))
    k_max = int(np.diff(sub_ptr).max())
    assert k_max > 1024  # the overlap makes the largest 1026: the block path is off
    pc = _pc_sym("additive", "float16")
    pc.setup(_torch_csr(A), decomposition=d)
    assert pc._msym is None
    # the same sizes without overlap reach exactly 1024
    d = P.Decomposition(perm=np.arange(n), owner_ptr=owner_ptr, sub_ptr=owner_ptr,
                        sub_idx=np.arange(n), solvers_per_coarse=np.array([4, 4]))
    for storage in ("float16", "bfloat16"):
        pc = _pc_sym("additive", storage)
        pc.setup(_torch_csr(A), decomposition=d)
        assert pc._msym is not None and pc.local.max_k == 1024
        R = np.random.default_rng(2).uniform(-1, 1, (n, 16))
        yl = pc._local_multi_into(torch.as_tensor(R, device=dev), None).cpu().numpy()
        O = SchwarzOracle(A, d.owner_ptr, d.sub_ptr, d.sub_idx, d.solvers_per_coarse,
                          kind="additive", inverses=pc.stored_local_inverses())
        ylo = np.stack([O.local(R[:, j]) for j in range(16)], axis=1)
        assert _colrel(yl, ylo).max() <= 1e-5, storage
