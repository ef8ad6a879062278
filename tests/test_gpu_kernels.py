"""GPU parity of the individual HIP kernels against the CPU oracle (through the C ABI)."""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sps
import torch

from helpers import bsr8
from oracle import kernels as K
from oracle.schwarz import SchwarzOracle

import dd_solvers as dds
from dd_solvers import _lib
from dd_solvers import problems as P
from dd_solvers.bsr import FastBSR
from dd_solvers.csr import csr_to_bsr
from dd_solvers.gemv import gemv_strided_batched

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")

DTYPE_PAIRS = [
    (torch.float16, torch.float32), (torch.float16, torch.float64),
    (torch.bfloat16, torch.float32), (torch.bfloat16, torch.float64),
    (torch.float32, torch.float32), (torch.float32, torch.float64),
    (torch.float64, torch.float64),
]


@pytest.mark.parametrize("mat_dtype,vec_dtype", DTYPE_PAIRS)
@pytest.mark.parametrize("k", [1, 2, 5, 10, 20, 50, 100, 200])
@pytest.mark.parametrize("use_shared_memory", [True, False])
@pytest.mark.parametrize("via", ["custom_op", "native_op"])
def test_gemv_strided_batched(mat_dtype, vec_dtype, k, use_shared_memory, via):
    # the 112 cases of src/dd_solvers/tests/test_gemv.py, checked against the oracle, through
    # the Python custom op (gemv.py) and through the native op that dd_solvers._C registers
    # (torch.ops.dd_solvers_gemv, the reference's gemv.cpp:55-57)
    n = 7
    g = torch.Generator().manual_seed(42)
    A = torch.randn((n, k, k), generator=g, dtype=torch.float32).to(mat_dtype)
    A = A.permute(0, 2, 1).contiguous().permute(0, 2, 1).to(dev)  # column-major
    b = torch.randn((n, k), generator=g, dtype=torch.float32).to(vec_dtype).to(dev)
    if via == "custom_op":
        out = gemv_strided_batched(A, b, use_shared_memory=use_shared_memory)
    else:
        out = torch.ops.dd_solvers_gemv.gemv_strided_batched(A, b, use_shared_memory)
        assert torch.equal(out, gemv_strided_batched(A, b, use_shared_memory=use_shared_memory))
    ref = K.gemv_strided_batched(A.to(torch.float64).cpu().numpy(), b.cpu().numpy())
    assert out.dtype == vec_dtype
    tol = 1e-12 if vec_dtype == torch.float64 else 1e-5
    assert np.allclose(out.cpu().numpy(), ref, rtol=tol, atol=tol)


def test_native_op_rejects_bad_inputs():
    op = torch.ops.dd_solvers_gemv.gemv_strided_batched
    A = torch.randn(3, 4, 4, dtype=torch.float64, device=dev)  # row-major
    with pytest.raises(RuntimeError, match="column-major"):
        op(A, torch.randn(3, 4, dtype=torch.float64, device=dev), False)
    Ac = A.transpose(1, 2).contiguous().transpose(1, 2)
    with pytest.raises(RuntimeError, match="Double-Float"):
        op(Ac, torch.randn(3, 4, dtype=torch.float32, device=dev), True)
    with pytest.raises(RuntimeError, match="compatible"):
        op(Ac, torch.randn(3, 5, dtype=torch.float64, device=dev), False)


def test_gemv_rejects_bad_inputs():
    A = torch.randn(3, 4, 4, dtype=torch.float64, device=dev)  # row-major
    with pytest.raises(RuntimeError, match="column-major"):
        gemv_strided_batched(A, torch.randn(3, 4, dtype=torch.float64, device=dev))
    Ac = A.transpose(1, 2).contiguous().transpose(1, 2)
    with pytest.raises(RuntimeError):
        gemv_strided_batched(Ac, torch.randn(3, 4, dtype=torch.float32, device=dev))
    with pytest.raises(RuntimeError, match="compatible"):
        gemv_strided_batched(Ac, torch.randn(3, 5, dtype=torch.float64, device=dev))


def _bsr_tensor(crow, col, vals, dtype):
    b = vals.shape[1]
    nbr = crow.shape[0] - 1
    return torch.sparse_bsr_tensor(torch.as_tensor(crow), torch.as_tensor(col),
                                   torch.as_tensor(vals, dtype=dtype), size=(nbr * b, nbr * b)).to(dev)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bsr_fixture(dtype):
    crow, col, vals = bsr8()  # test_bsr.py:9-138 (b = 3: the lane-per-block kernel)
    M = _bsr_tensor(crow, col, vals, dtype)
    x = torch.randn(M.shape[0], dtype=dtype, device=dev)
    y = FastBSR(M) @ x
    ref = K.bsrmv(crow, col, vals.astype(np.float64 if dtype == torch.float64 else np.float32),
                  x.cpu().numpy().astype(np.float64))
    assert np.allclose(y.cpu().numpy(), ref, atol=1e-12 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("b", [2, 3, 4, 6, 8, 10, 15, 20, 21, 28])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bsr_ragged_rows(b, dtype):
    """Block rows of 0 to 9 blocks (empty rows included) through every streaming shape:
    the tail steps of a lane group and groups sharing a wavefront; with x . y partials."""
    rng = np.random.default_rng(100 + b)
    nbr = 3001
    cnt = rng.integers(0, 10, size=nbr)
    cnt[::97] = 0
    crow = np.concatenate([[0], np.cumsum(cnt)])
    col = rng.integers(0, nbr, size=int(crow[-1]))
    vals = rng.standard_normal((int(crow[-1]), b, b))
    M = _bsr_tensor(crow, col, vals, dtype)
    op = FastBSR(M)
    x = torch.as_tensor(rng.standard_normal(nbr * b), dtype=dtype, device=dev)
    y = torch.empty_like(x)
    partials = torch.zeros(_lib.MAX_PARTIALS, dtype=torch.float64, device=dev)
    n = op._enqueue_matvec(x, y, partials, None)
    ref = K.bsrmv(crow, col, vals.astype(np.float32 if dtype == torch.float32 else np.float64),
                  x.cpu().numpy().astype(np.float64))
    tol = 1e-12 if dtype == torch.float64 else 1e-4
    assert np.allclose(y.cpu().numpy(), ref, atol=tol * max(1.0, np.abs(ref).max()))
    assert np.isclose(partials[:n].sum().item(), float(ref @ x.cpu().numpy().astype(np.float64)),
                      rtol=1e-10 if dtype == torch.float64 else 1e-4)


@pytest.mark.parametrize("b", [4, 6, 10, 20, 28])
def test_bsr_values_off_the_pair_alignment(b):
    """Values starting one element past a 16-byte boundary (a caller's view into a larger
    buffer) take the one-element kernels, with the same result."""
    rng = np.random.default_rng(b)
    nbr = 500
    crow = np.arange(0, 4 * nbr + 1, 4)
    col = rng.integers(0, nbr, size=4 * nbr)
    vals = rng.standard_normal((4 * nbr, b, b))
    buf = torch.empty(vals.size + 1, dtype=torch.float64, device=dev)
    buf[1:] = torch.as_tensor(np.ascontiguousarray(vals.transpose(0, 2, 1)).reshape(-1), device=dev)
    x = torch.as_tensor(rng.standard_normal(nbr * b), device=dev)
    y = torch.empty_like(x)
    crow_t = torch.as_tensor(crow, dtype=torch.int64, device=dev)
    col_t = torch.as_tensor(col, dtype=torch.int32, device=dev)
    assert buf[1:].data_ptr() % 16 == 8
    _lib.call_partials("dds_bsrmv", _lib.F64, _lib.I64, nbr, b, crow_t.data_ptr(), col_t.data_ptr(),
                       buf[1:].data_ptr(), x.data_ptr(), y.data_ptr(), None, _lib.PARTIALS_OUT,
                       None, _lib.stream())
    ref = K.bsrmv(crow, col, vals, x.cpu().numpy())
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("b", [1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 15, 20, 21, 28])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bsr_random(b, dtype):
    # test_bsr.py:161-193: 1000 block rows, 3 random blocks per row (duplicates allowed)
    rng = np.random.default_rng(b)
    nbr = 1000
    crow = np.arange(0, 3 * nbr + 1, 3)
    col = rng.integers(0, nbr, size=3 * nbr)
    vals = rng.standard_normal((3 * nbr, b, b))
    M = _bsr_tensor(crow, col, vals, dtype)
    x = torch.as_tensor(rng.standard_normal(nbr * b), dtype=dtype, device=dev)
    y = FastBSR(M) @ x
    ref = K.bsrmv(crow, col, vals.astype(np.float32 if dtype == torch.float32 else np.float64),
                  x.cpu().numpy().astype(np.float64))
    assert np.allclose(y.cpu().numpy(), ref, atol=1e-12 if dtype == torch.float64 else 1e-4)


@pytest.mark.parametrize("b", [2, 4, 8])
def test_csr_to_bsr_then_spmv_q1(b):
    A = P.assemble_q1_3d(16, P.lognormal_coefficients(16 ** 3, 1).reshape(16, 16, 16))
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                 torch.as_tensor(A.data), A.shape).to(dev)
    B = csr_to_bsr(At, b)
    assert B.values().stride() == (b * b, 1, b)
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    xp = np.zeros(B.shape[0])
    xp[:A.shape[0]] = x
    y = FastBSR(B) @ torch.as_tensor(xp, device=dev)
    assert np.allclose(y.cpu().numpy()[:A.shape[0]], A @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind", ["q1_3d", "p1_2d"])
def test_csrmv(kind):
    pb = P.make_problem(kind, 24 if kind == "q1_3d" else 64, box=4, overlap=0, lognormal=True)
    A = pb.A
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                 torch.as_tensor(A.data), A.shape).to(dev)
    op = dds.solvers._CSROperator(At)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    y = op @ torch.as_tensor(x, device=dev)
    assert np.allclose(y.cpu().numpy(), A @ x, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------------------
# coarse solver
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("value_type", [torch.float32, torch.float64])
@pytest.mark.parametrize("index_type", [torch.int32, torch.int64])
@pytest.mark.parametrize("mode", ["trsv", "inverse"])
def test_dense_cholesky_known_answers(value_type, index_type, mode):
    # src/dd_solvers/tests/test_cudss.py:34-58
    A = torch.sparse_csr_tensor(torch.tensor([0, 2, 3, 5], dtype=index_type),
                                torch.tensor([0, 2, 1, 0, 2], dtype=index_type),
                                torch.tensor([2, 2, 2, 2, 3], dtype=value_type), size=(3, 3)).to(dev)
    s = dds.DenseCholesky(mode=mode)
    s.setup(A)
    for b, x in (([2, -2, 2], [1, -1, 0]), ([4, 4, 5], [1, 2, 1])):
        xh, _ = s.solve(torch.tensor(b, dtype=value_type, device=dev))
        assert xh.dtype == value_type
        assert torch.allclose(xh.cpu(), torch.tensor(x, dtype=value_type), atol=1e-6)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 300, 1000, 2100])
@pytest.mark.parametrize("mode", ["trsv", "inverse"])
def test_dense_cholesky_random(n, mode):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T / n + np.eye(n)
    b = rng.standard_normal(n)
    s = dds.DenseCholesky(mode=mode)
    s.setup(torch.as_tensor(A, device=dev).to_sparse_csr())
    L = np.tril(s.L.cpu().numpy())
    assert np.allclose(L, K.cholesky_lower(A), rtol=1e-11, atol=1e-11)
    x, _ = s.solve(torch.as_tensor(b, device=dev))
    xr = sla.cho_solve(sla.cho_factor(A, lower=True), b)
    assert np.linalg.norm(x.cpu().numpy() - xr) <= 1e-11 * np.linalg.norm(xr)


def test_dense_cholesky_not_spd():
    A = torch.tensor([[1.0, 2.0], [2.0, 1.0]], dtype=torch.float64, device=dev)
    with pytest.raises(np.linalg.LinAlgError):
        dds.DenseCholesky().setup(A)


# ---------------------------------------------------------------------------------------
# local inverses (ragged, overlapping)
# ---------------------------------------------------------------------------------------

def _torch_csr(A):
    return torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                   torch.as_tensor(A.data), A.shape).to(dev)


@pytest.mark.parametrize("storage", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("overlap", [0, 1])
@pytest.mark.parametrize("layout,waves", [("full", 1), ("full", 4), ("full", 8), ("sym", None),
                                          ("sym_mfma", None)])
@pytest.mark.parametrize("N,box", [(14, 4), (18, 8)])
def test_local_apply_matches_oracle(storage, overlap, layout, waves, N, box):
    # (18, 8): subdomains of 8^3 + overlap -> k up to 1000, several 64-row tiles
    if layout == "sym_mfma" and storage not in (torch.float16, torch.bfloat16):
        pytest.skip("the matrix-core layout stores FP16 or BF16")
    print(f"local apply: storage {storage} overlap {overlap} layout {layout} N {N} box {box}", flush=True)
    pb = P.make_problem("q1_3d", N, box=box, overlap=overlap, seed=2, lognormal=True)
    d = pb.decomposition
    scaled = storage == torch.float16
    acc = torch.float32 if layout == "sym_mfma" else None
    M = dds.AdditiveSchwarz(torch.float64, dds.Inv(storage, scaled=scaled, layout=layout,
                                                   accumulate=acc), dds.DenseCholesky())
    M.setup(_torch_csr(pb.A), decomposition=d, waves=waves)
    assert M.local.layout == layout
    sname = {torch.float64: "float64", torch.float32: "float32", torch.float16: "float16",
             torch.bfloat16: "bfloat16"}[storage]
    r = np.random.default_rng(0).standard_normal(pb.A.shape[0])
    y = M._local_only(torch.as_tensor(r, device=dev)).cpu().numpy()
    # 1. the kernel's arithmetic: the exact FP64 product of the inverses the device stores.
    # FP64 accumulation leaves rounding only; the matrix-core apply keeps ~22 (FP16) or ~24
    # (BF16) bits of x / sigma_s and accumulates in FP32.
    Oe = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, inverses=M.stored_local_inverses())
    ye = Oe.local(r)
    rel_exact = np.linalg.norm(y - ye) / np.linalg.norm(ye)
    assert rel_exact < (2e-6 if layout == "sym_mfma" else 1e-13), rel_exact
    # 2. the whole set-up: inversion and storage rounding against the oracle's own inverses
    O = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx, storage=sname, scaled=scaled)
    yo = O.local(r)
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    tol = {torch.float64: 1e-12, torch.float32: 1e-6, torch.float16: 2e-3, torch.bfloat16: 1e-2}[storage]
    assert rel < tol, rel


@pytest.mark.parametrize("storage", [torch.float16, torch.bfloat16])
@pytest.mark.parametrize("group", [1, 2, 3, "all"])
@pytest.mark.parametrize("N,box,overlap", [(18, 8, 1), (14, 4, 0), (20, 8, 0)])
def test_sym_mfma_pair_groups(monkeypatch, storage, group, N, box, overlap):
    """Items of G tile-row pairs (partial vectors summed by the combine), and one item per
    subdomain writing the result directly: the exact product of the stored inverses; the
    pack and the apply launched apart give the same bits as the entry point that runs both."""
    pb = P.make_problem("q1_3d", N, box=box, overlap=overlap, seed=3, lognormal=True)
    d = pb.decomposition
    sizes = np.diff(d.sub_ptr)
    P_max = -(-(-(-int(sizes.max()) // 64)) // 2)
    G = P_max if group == "all" else group
    print(f"sym_mfma groups: storage {storage} G {G} (P_max {P_max}) N {N} box {box} overlap {overlap}",
          flush=True)
    monkeypatch.setattr(dds.solvers, "_symm_group", lambda sizes, n_cu: G)
    M = dds.AdditiveSchwarz(torch.float64, dds.Inv(storage, scaled=storage == torch.float16,
                                                   accumulate=torch.float32, layout="sym_mfma"),
                            dds.DenseCholesky())
    M.setup(_torch_csr(pb.A), decomposition=d)
    L = M.local
    assert L.layout == "sym_mfma" and L.group == G and L.direct == (G >= P_max)
    r = np.random.default_rng(1).standard_normal(pb.A.shape[0])
    rt = torch.as_tensor(r, device=dev)
    y = M._local_only(rt).cpu().numpy()
    ye = SchwarzOracle(pb.A, d.owner_ptr, d.sub_ptr, d.sub_idx,
                       inverses=M.stored_local_inverses()).local(r)
    rel = np.linalg.norm(y - ye) / np.linalg.norm(ye)
    assert rel < 2e-6, rel
    # the two halves: the pack, then the apply from it (and again from the same pack)
    Yst = M.Y if M.Y is not None else torch.empty_like(rt)
    Y1 = torch.full_like(Yst, float("nan"))
    M._enqueue_sym_apply(rt, None, Y1)
    if not L.direct:
        M._enqueue_sym_combine(Y1, None)
    M._enqueue_sym_pack(rt, None)
    for _ in range(2):
        Y2 = torch.full_like(Yst, float("nan"))
        M._enqueue_sym_apply_packed(None, Y2)
        if not L.direct:
            M._enqueue_sym_combine(Y2, None)
        torch.cuda.synchronize()
        assert not torch.isnan(Y2).any()
        assert torch.equal(Y1, Y2)


@pytest.mark.parametrize("pi", [torch.float64, torch.float32])
@pytest.mark.parametrize("group", [1, "all"])
@pytest.mark.parametrize("N,box,overlap", [(18, 8, 1), (20, 8, 0)])
def test_sym_mfma_fused_pack_same_bits(monkeypatch, pi, group, N, box, overlap):
    """FP16 tiles: the apply that packs x in its own LDS gives the bits of the pack kernel
    followed by the apply, for FP64 and FP32 vectors, with and without overlap (gathered x)."""
    pb = P.make_problem("q1_3d", N, box=box, overlap=overlap, seed=5, lognormal=True)
    d = pb.decomposition
    sizes = np.diff(d.sub_ptr)
    P_max = -(-(-(-int(sizes.max()) // 64)) // 2)
    G = P_max if group == "all" else group
    print(f"sym_mfma fused pack: pi {pi} G {G} N {N} box {box} overlap {overlap}", flush=True)
    monkeypatch.setattr(dds.solvers, "_symm_group", lambda sizes, n_cu: G)
    M = dds.AdditiveSchwarz(pi, dds.Inv(torch.float16, scaled=True, accumulate=torch.float32,
                                        layout="sym_mfma"), dds.DenseCholesky())
    M.setup(_torch_csr(pb.A), decomposition=d)
    L = M.local
    assert L.layout == "sym_mfma" and L.direct == (G >= P_max)
    rt = torch.as_tensor(np.random.default_rng(2).standard_normal(pb.A.shape[0]), device=dev).to(pi)
    Yst = M.Y if M.Y is not None else torch.empty_like(rt)
    out = []
    for fused in (True, False):
        M.fused_pack = fused
        Y = torch.full_like(Yst, float("nan"))
        M._enqueue_sym_apply(rt, None, Y)
        if not L.direct:
            M._enqueue_sym_combine(Y, None)
        torch.cuda.synchronize()
        assert not torch.isnan(Y).any()
        out.append(Y)
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("kind,N", [("q1_3d", 24), ("p1_2d", 64), ("q1_3d", 5)])
@pytest.mark.parametrize("fmt", ["sell", "csr"])
def test_spmv_formats_with_dot(kind, N, fmt):
    pb = P.make_problem(kind, N, box=4, overlap=0, lognormal=True)
    A = pb.A
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                 torch.as_tensor(A.data), A.shape).to(dev)
    op = dds.solvers._SELLOperator(At) if fmt == "sell" else dds.solvers._CSROperator(At)
    x = np.random.default_rng(1).standard_normal(A.shape[0])
    xt = torch.as_tensor(x, device=dev)
    y = torch.empty_like(xt)
    part = torch.zeros(2048, dtype=torch.float64, device=dev)
    cnt = op._enqueue_matvec(xt, y, part, None)
    torch.cuda.synchronize()
    assert np.allclose(y.cpu().numpy(), A @ x, rtol=1e-13, atol=1e-13)
    assert np.isclose(part[:cnt].sum().item(), x @ (A @ x), rtol=1e-12)


@pytest.mark.parametrize("lanes", [None, 64, 128, 256, 512, 1024])
@pytest.mark.parametrize("value_type,index_type", [(torch.float64, torch.int64),
                                                   (torch.float64, torch.int32),
                                                   (torch.float32, torch.int64)])
def test_csrmv_long_rows(lanes, value_type, index_type):
    """Rows of 0 .. 3000 entries (the shape of R_C A, plus empty rows), every lane setting
    including the workgroup-per-row kernel, with the x . y epilogue on a square matrix."""
    rng = np.random.default_rng(7)
    n = 1500
    lens = rng.integers(0, 3000, n)
    lens[::97] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, lens.sum())
    A = sps.csr_matrix((rng.standard_normal(lens.sum()), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    if value_type == torch.float32:  # the values and x the kernel sees, in float64
        A.data = A.data.astype(np.float32).astype(np.float64)
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr).to(index_type),
                                 torch.as_tensor(A.indices.astype(np.int32)),
                                 torch.as_tensor(A.data).to(value_type), A.shape).to(dev)
    op = dds.solvers._CSROperator(At, lanes=lanes)
    if lanes is None:
        assert op.lanes == 256  # average row length > 128
    xt = torch.as_tensor(rng.standard_normal(n), device=dev).to(value_type)
    x = xt.double().cpu().numpy()
    y = torch.empty_like(xt)
    part = torch.zeros(2048, dtype=torch.float64, device=dev)
    cnt = op._enqueue_matvec(xt, y, part, None)
    torch.cuda.synchronize()
    ref = A @ x  # double accumulation, as in the kernels
    tol = 1e-13 if value_type == torch.float64 else 1e-6
    scale = np.abs(A) @ np.abs(x)
    assert np.all(np.abs(y.double().cpu().numpy() - ref) <= tol * scale + 1e-300)
    assert abs(part[:cnt].sum().item() - ref @ x) <= 1e-12 * (scale @ np.abs(x))
    # deterministic: the same launch gives the same bits
    y2 = torch.empty_like(xt)
    op._enqueue_matvec(xt, y2, None, None)
    torch.cuda.synchronize()
    assert torch.equal(y, y2)


def test_csrmv_tall_matrix():
    """Rows > columns (the Q1 assembly operator): y = G k must not touch k past its end."""
    N = 16
    G = P.q1_assembly_operator(N)
    Gt = torch.sparse_csr_tensor(torch.as_tensor(G.indptr.astype(np.int64)),
                                 torch.as_tensor(G.indices.astype(np.int32)),
                                 torch.as_tensor(G.data), G.shape).to(dev)
    op = dds.solvers._CSROperator(Gt)
    k = np.random.default_rng(0).random(N ** 3)
    y = op @ torch.as_tensor(k, device=dev)
    torch.cuda.synchronize()
    assert np.allclose(y.cpu().numpy(), G @ k, rtol=1e-14, atol=1e-15)
    with pytest.raises(ValueError):
        op._enqueue_matvec(torch.as_tensor(k, device=dev), y,
                           torch.zeros(2048, dtype=torch.float64, device=dev), None)


# ---------------------------------------------------------------------------------------
# batched SPD inverse (local matrices at setup)
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 2e-3)])
def test_spd_inverse_batched_matches_torch(dtype, tol):
    from dd_solvers import _lib
    rng = np.random.default_rng(5)
    ks = [1, 5, 31, 32, 33, 64, 65, 100, 130, 257]
    kmax = max(ks)
    A = np.zeros((len(ks), kmax, kmax))
    for b, k in enumerate(ks):
        B = rng.standard_normal((k, k))
        A[b, :k, :k] = B @ B.T + k * np.eye(k)
        A[b, k:, k:] = np.eye(kmax - k) * 7.0  # padding outside k x k must stay as it is
    At = torch.as_tensor(A, dtype=dtype, device=dev).contiguous()
    orig = At.clone()
    info = torch.full((len(ks),), -1, dtype=torch.int32, device=dev)
    kt = torch.as_tensor(ks, dtype=torch.int32, device=dev)
    _lib.call("dds_spd_inverse_batched", _lib.dtype_code(dtype), len(ks), kt.data_ptr(), kmax,
              kmax * kmax, At.data_ptr(), info.data_ptr(), _lib.stream())
    assert info.cpu().tolist() == [0] * len(ks)
    for b, k in enumerate(ks):
        ref = np.linalg.inv(A[b, :k, :k])
        got = At[b, :k, :k].double().cpu().numpy()
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), (k, np.abs(got - ref).max())
        assert torch.equal(At[b, k:, :], orig[b, k:, :]) and torch.equal(At[b, :k, k:], orig[b, :k, k:])


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-3)])
def test_inverse_batched_pivoting(dtype, tol):
    """dds_inverse_batched (partial pivoting) against numpy on nonsymmetric, zero-diagonal
    and indefinite matrices of several sizes; the padding outside each k x k stays."""
    rng = np.random.default_rng(11)
    ks = [1, 2, 7, 33, 64, 65, 130, 300]
    kmax = max(ks)
    mats = []
    for k in ks:
        B = rng.standard_normal((k, k))
        Z = B.copy()
        np.fill_diagonal(Z, 0.0)
        mats += [B + 3 * np.eye(k), Z if k > 1 else B, B + B.T]
    A = np.zeros((len(mats), kmax, kmax))
    kk = []
    for b, M in enumerate(mats):
        k = M.shape[0]
        kk.append(k)
        A[b, :k, :k] = M
        A[b, k:, k:] = np.eye(kmax - k) * 7.0
    At = torch.as_tensor(A, dtype=dtype, device=dev).contiguous()
    orig = At.clone()
    info = torch.zeros(len(mats), dtype=torch.int32, device=dev)
    kt = torch.as_tensor(kk, dtype=torch.int32, device=dev)
    _lib.call("dds_inverse_batched", _lib.dtype_code(dtype), len(mats), kt.data_ptr(), kmax,
              kmax * kmax, At.data_ptr(), info.data_ptr(), _lib.stream())
    assert info.cpu().tolist() == [0] * len(mats)
    for b, M in enumerate(mats):
        k = M.shape[0]
        ref = np.linalg.inv(M)
        got = At[b, :k, :k].double().cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= tol * max(1.0, np.linalg.cond(M) / 100), (b, k, err)
        assert torch.equal(At[b, k:, :], orig[b, k:, :]) and torch.equal(At[b, :k, k:], orig[b, :k, k:])


def test_inverse_batched_singular_and_inv_setup():
    """A singular matrix raises torch.linalg.LinAlgError (as torch.linalg.inv does); Inv.setup
    inverts any nonsingular batch and Inv.solve applies it."""
    from dd_solvers.solvers import inverse_batched
    rng = np.random.default_rng(12)
    A = rng.standard_normal((3, 6, 6))
    A[1, :, 2] = 0.0
    with pytest.raises(torch.linalg.LinAlgError, match="matrix 1 is singular"):
        inverse_batched(torch.as_tensor(A, device=dev))
    A[1] = rng.standard_normal((6, 6))
    np.fill_diagonal(A[2], 0.0)
    inv = dds.Inv(torch.float64)
    inv.setup(torch.as_tensor(A, device=dev))
    rhs = rng.standard_normal((3, 6))
    out = inv.solve(torch.as_tensor(rhs, device=dev)).cpu().numpy()
    assert np.allclose(out, np.linalg.solve(A, rhs[..., None])[..., 0], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("layout", ["full", "sym"])
def test_schwarz_setup_inverts_indefinite_local_matrices(layout):
    """Local matrices that are not SPD (the SPD Gauss-Jordan flags them) are inverted again
    with pivoting, end to end through the preconditioner setup: A - c v v^T with
    v = e_i - e_j for two DoFs of one coarse cell leaves the coarse matrix as it is and makes
    every local matrix holding i or j indefinite."""
    import scipy.sparse as sp
    pb = P.make_problem("q1_3d", 12, box=4, overlap=1, seed=4, lognormal=True)
    d = pb.decomposition
    A = pb.A.tocsr()
    i, j = int(d.owner_ptr[0]), int(d.owner_ptr[0]) + 1
    c = 10.0 * A.diagonal().max()
    vvT = sp.coo_matrix(([1.0, -1.0, -1.0, 1.0], ([i, i, j, j], [i, j, i, j])), shape=A.shape)
    Ap = (A - c * vvT).tocsr()
    Ap.sort_indices()
    M = dds.AdditiveSchwarz(torch.float64, dds.Inv(torch.float64, layout=layout), dds.DenseCholesky())
    M.setup(_torch_csr(Ap), decomposition=d)
    assert M.local.layout == layout
    got = M.stored_local_inverses()
    sub_ptr, sub_idx = d.sub_ptr, d.sub_idx
    n_indef = 0
    for s_ in range(len(sub_ptr) - 1):
        idx = sub_idx[sub_ptr[s_]:sub_ptr[s_ + 1]] if sub_idx is not None else np.arange(sub_ptr[s_], sub_ptr[s_ + 1])
        Al = Ap[idx][:, idx].toarray()
        n_indef += int(np.linalg.eigvalsh(Al).min() < 0)
        ref = np.linalg.inv(Al)
        err = np.linalg.norm(got[s_] - ref) / np.linalg.norm(ref)
        assert err < 1e-13 * np.linalg.cond(Al), (s_, err)
    assert n_indef >= 1


def test_spd_inverse_batched_flags_indefinite():
    from dd_solvers import _lib
    A = torch.tensor([[[2.0, 0.0], [0.0, -1.0]], [[4.0, 1.0], [1.0, 3.0]]], dtype=torch.float64, device=dev)
    info = torch.zeros(2, dtype=torch.int32, device=dev)
    kt = torch.tensor([2, 2], dtype=torch.int32, device=dev)
    _lib.call("dds_spd_inverse_batched", _lib.F64, 2, kt.data_ptr(), 2, 4, A.data_ptr(),
              info.data_ptr(), _lib.stream())
    assert info.cpu().tolist() == [1, 0]
    assert np.allclose(A[1].cpu().numpy(), np.linalg.inv([[4.0, 1.0], [1.0, 3.0]]), rtol=1e-14)


def test_spd_inverse_rejects_bad_arguments():
    from dd_solvers import _lib
    with pytest.raises(_lib.HipError):
        _lib.call("dds_spd_inverse_batched", _lib.F16, 1, 1, 2, 4, 1, None, None)


def test_stream_kernels():
    from dd_solvers import _lib
    n = 1 << 16
    a = torch.rand(n, dtype=torch.float64, device=dev)
    b = torch.rand(n, dtype=torch.float64, device=dev)
    y = torch.empty_like(a)
    s = _lib.stream()
    _lib.call("dds_stream", 1, n, a.data_ptr(), None, 0.0, y.data_ptr(), 7, s)
    assert torch.equal(y, a)
    _lib.call("dds_stream", 2, n, a.data_ptr(), b.data_ptr(), 0.5, y.data_ptr(), 3, s)
    assert torch.allclose(y, a + 0.5 * b, rtol=1e-15, atol=0)
    part = torch.empty(5, dtype=torch.float64, device=dev)
    _lib.call("dds_stream", 0, n, a.data_ptr(), None, 0.0, part.data_ptr(), 5, s)
    assert abs(part.sum().item() - a.sum().item()) <= 1e-9 * a.sum().item()
    with pytest.raises(_lib.HipError):
        _lib.call("dds_stream", 1, n + 1, a.data_ptr(), None, 0.0, y.data_ptr(), 4, s)


def test_sell_mixed_compact_and_wide_slices():
    """Rows of the first half each have one scattered far column (their slots span more
    than 65535 columns, so those slices keep int32 columns), the rest is banded (16-bit
    offsets); with the fused x . y partials."""
    import scipy.sparse as sps
    from dd_solvers import _lib
    from dd_solvers.solvers import _SELLOperator
    n = 150_000
    rng = np.random.default_rng(2)
    rows, cols = [], []
    for off in (-1, 0, 1):
        r = np.arange(max(0, -off), min(n, n - off))
        rows.append(r)
        cols.append(r + off)
    far = np.arange(n // 2)
    rows.append(far)
    cols.append((far * 7919 + 12345) % n)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    A = sps.csr_matrix((rng.standard_normal(r.shape[0]), (r, c)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                 torch.as_tensor(A.data), A.shape).to(dev)
    op = _SELLOperator(At)
    assert 0.3 < op.wide_fraction < 0.7
    x = torch.as_tensor(rng.standard_normal(n), device=dev)
    y = torch.empty_like(x)
    partials = torch.zeros(_lib.MAX_PARTIALS, dtype=torch.float64, device=dev)
    cnt = op._enqueue_matvec(x, y, partials, None)
    ref = A @ x.cpu().numpy()
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    assert np.isclose(partials[:cnt].sum().item(), float(ref @ x.cpu().numpy()), rtol=1e-12)


@pytest.mark.parametrize("lanes", [0, 8, 16, 32, 64])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_restrict_all_lane_counts(lanes, dtype):
    """R_C x: per-cell sums over cells of 0 to 700 unknowns (empty cells included), every
    launch shape, against numpy."""
    rng = np.random.default_rng(lanes)
    sizes = np.concatenate([rng.integers(0, 12, 3000), rng.integers(40, 700, 50), [0, 1, 3]])
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    x = rng.standard_normal(int(ptr[-1]))
    xt = torch.as_tensor(x, device=dev).to(dtype)
    xc = torch.full((sizes.shape[0],), 7.0, dtype=dtype, device=dev)
    _lib.call("dds_restrict", _lib.dtype_code(dtype), sizes.shape[0],
              torch.as_tensor(ptr, device=dev).data_ptr(), xt.data_ptr(), xc.data_ptr(), lanes,
              None, _lib.stream())
    ref = np.array([x.astype(np.float32 if dtype == torch.float32 else np.float64)[a:b].sum()
                    for a, b in zip(ptr[:-1], ptr[1:])])
    got = xc.double().cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 1e-4
    assert np.allclose(got, ref, rtol=tol, atol=tol)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 127, 128, 129, 1000, 4097])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("offset", [0, 1, 2, 3])
def test_dense_symv_all_widths(n, dtype, offset):
    """y = S x for a dense row-major S: rows on 16-byte boundaries (chunked loads and the
    element tail) and rows off them by every element count (offset start and leading
    dimension n + offset: the head elements up to the boundary, then chunks with x read
    element-wise)."""
    rng = np.random.default_rng(n)
    lda = n + offset
    S = rng.standard_normal((n, lda))
    x = rng.standard_normal(n)
    St = torch.as_tensor(np.concatenate([np.zeros(offset), S.reshape(-1)]), device=dev).to(dtype)
    xt = torch.as_tensor(x, device=dev).to(dtype)
    yt = torch.empty(n, dtype=dtype, device=dev)
    _lib.call("dds_dense_symv", _lib.dtype_code(dtype), n, St[offset:].data_ptr(), lda,
              xt.data_ptr(), yt.data_ptr(), None, _lib.stream())
    ref = S[:, :n] @ x
    tol = 1e-12 if dtype == torch.float64 else 2e-4
    assert np.allclose(yt.double().cpu().numpy(), ref, rtol=tol, atol=tol * np.sqrt(n))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_segment_sum_small_and_empty_runs(dtype):
    """dds_segment_sum adds each run in index order, in the values' dtype; empty runs give 0."""
    from dd_solvers.csr import segment_sum
    rng = np.random.default_rng(5)
    lengths = rng.integers(0, 9, size=300)
    vals = rng.standard_normal(int(lengths.sum())).astype(np.float64 if dtype == torch.float64 else np.float32)
    got = segment_sum(torch.as_tensor(vals, device=dev), torch.as_tensor(lengths, device=dev))
    want, k = [], 0
    for ln in lengths:
        acc = vals.dtype.type(0)
        for v in vals[k:k + ln]:
            acc = vals.dtype.type(acc + v)
        want.append(acc)
        k += ln
    assert got.dtype == dtype
    assert np.array_equal(got.cpu().numpy(), np.array(want, dtype=vals.dtype))


def test_segment_sum_beyond_2_pow_24_segments():
    """More than 2^24 runs (the merge of R_C A for the reference's p = 3 S9 / S9 / S9 runs has
    21 million): every run's sum equals the in-order sum (built on the host position by
    position: the j-th entry of every run is added at step j)."""
    from dd_solvers.csr import segment_sum
    rng = np.random.default_rng(6)
    n = (1 << 24) + 4096
    lengths = rng.integers(1, 8, size=n)
    vals = rng.standard_normal(int(lengths.sum()))
    got = segment_sum(torch.as_tensor(vals, device=dev), torch.as_tensor(lengths, device=dev))
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    want = np.zeros(n)
    for j in range(int(lengths.max())):
        has = lengths > j
        want[has] = want[has] + vals[starts[has] + j]
    assert np.array_equal(got.cpu().numpy(), want)


def test_segment_sum_compensated_keeps_cancelled_terms():
    """Compensated runs keep what a plain in-order sum rounds away: 1e16 + 1 - 1e16 is 0
    plainly and 1 compensated; on random runs the compensated sums equal the exact (math.fsum)
    sums to one rounding."""
    import math
    from dd_solvers.csr import segment_sum
    v = torch.tensor([1e16, 1.0, -1e16, 3.0, 1e-20, -3.0], dtype=torch.float64, device=dev)
    ln = torch.tensor([3, 3], device=dev)
    assert segment_sum(v, ln).tolist() == [0.0, 0.0]
    assert segment_sum(v, ln, compensated=True).tolist() == [1.0, 1e-20]
    rng = np.random.default_rng(8)
    lengths = rng.integers(1, 2000, size=200)
    vals = rng.standard_normal(int(lengths.sum())) * 10.0 ** rng.integers(-8, 9, size=int(lengths.sum()))
    got = segment_sum(torch.as_tensor(vals, device=dev), torch.as_tensor(lengths, device=dev),
                      compensated=True).cpu().numpy()
    k = 0
    for g, ln_ in zip(got, lengths):
        exact = math.fsum(vals[k:k + ln_])
        assert abs(g - exact) <= 4.5e-16 * abs(exact) + 1e-300
        k += ln_
