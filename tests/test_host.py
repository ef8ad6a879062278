"""CPU-side checks: the C-ABI library loads and exports every declared symbol, host-side
argument validation, the setup utilities (on CPU tensors), and the problem generators."""

import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from helpers import dg24

import dd_solvers as dds
from dd_solvers import _lib
from dd_solvers import problems as P
from dd_solvers.csr import asm_permutation, csr_to_bsr, gather_csr, inverse_permutation, \
    permute_csr_tensor, sort_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(dds_[a-z0-9_]+)\s*\(", txt))
    return names


def test_native_op_registered_with_reference_schema():
    """dd_solvers._C registers dd_solvers_gemv::gemv_strided_batched with the schema the
    reference's TORCH_LIBRARY infers from its function (gemv.cpp:9-12, 55-57); its checks
    run before anything touches a device."""
    import dd_solvers._C  # noqa: F401
    op = torch.ops.dd_solvers_gemv.gemv_strided_batched
    schema = str(op.default._schema)
    assert schema == "dd_solvers_gemv::gemv_strided_batched(Tensor _0, Tensor _1, bool _2) -> Tensor _0"
    with pytest.raises(RuntimeError, match="compatible dimensions"):
        op(torch.zeros(2, 3, 3), torch.zeros(2, 4), False)
    with pytest.raises(RuntimeError, match="column-major"):
        op(torch.zeros(2, 3, 3), torch.zeros(2, 3), False)
    with pytest.raises(RuntimeError, match="same GPU"):
        op(torch.zeros(2, 3, 3).transpose(1, 2), torch.zeros(2, 3), False)


def test_library_exports_every_declared_symbol():
    declared = _declared_symbols()
    assert "dds_local_apply" in declared and "dds_chol_factor" in declared
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) >= declared


def _declared_arity():
    """Number of parameters of every dds_* prototype in include/*.h."""
    out = {}
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for name, params in re.findall(r"\b(dds_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
                params = params.strip()
                out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_ctypes_signatures_match_the_header():
    """The ctypes binding declares as many arguments as the C prototype of each entry point
    (an argument missing from the binding would be passed with the default int conversion,
    truncating a 64-bit pointer)."""
    arity = _declared_arity()
    assert arity["dds_nd_factor"] == 23 and arity["dds_last_error"] == 0
    for name in _lib.EXPORTED:
        fn = getattr(_lib._lib, name)
        assert fn.argtypes is not None, name
        assert len(fn.argtypes) == arity[name], (name, len(fn.argtypes), arity[name])


def test_abi_version():
    assert _lib._lib.dds_abi_version() == _lib.ABI_VERSION


def test_host_validation_without_gpu():
    # Arguments are validated before any launch, so these return errors with no device.
    with pytest.raises(_lib.HipError, match="waves"):
        _lib.call("dds_local_apply", _lib.F64, _lib.F64, _lib.F64, 1, 1, 1, None, 8, 1, 1,
                  None, 1, 1, 1, 3, None, None)
    with pytest.raises(_lib.HipError, match="Double-Float"):
        _lib.call("dds_gemv_strided_batched", _lib.F64, _lib.F32, 1, 1, 1, 2, 2, 0, None)
    with pytest.raises(_lib.HipError, match="lanes"):
        _lib.call("dds_csrmv", _lib.F64, _lib.I64, 4, 1, 1, 1, 1, 1, 3, None, None, None, None)
    with pytest.raises(_lib.HipError, match="lanes"):
        _lib.call("dds_csrmv", _lib.F64, _lib.I64, 4, 1, 1, 1, 1, 1, 2048, None, None, None, None)
    with pytest.raises(_lib.HipError, match="bad op"):
        _lib.call("dds_pcg_finalize", 9, 1, 1, 1, 1, 1, 1, None)
    with pytest.raises(_lib.HipError, match="exceeds"):
        _lib.call("dds_chol_solve", _lib.F64, 7000, 1, 7000, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="unsupported"):
        _lib.call("dds_cast", _lib.F16, _lib.F64, 4, 1, 1, None, None)
    # entry points of the batched, setup and measurement paths
    with pytest.raises(_lib.HipError, match="dtype must be F64 or F32"):
        _lib.call("dds_spd_inverse_batched", _lib.F16, 2, 1, 4, 16, 1, None, None)
    with pytest.raises(_lib.HipError, match="ld/stride"):
        _lib.call("dds_spd_inverse_batched", _lib.F64, 2, 1, 4, 2, 1, None, None)
    with pytest.raises(_lib.HipError, match="n must be even"):
        _lib.call("dds_stream", 1, 7, 1, None, 0.0, 1, 8, None)
    with pytest.raises(_lib.HipError, match="grid"):
        _lib.call("dds_stream", 1, 8, 1, None, 0.0, 1, _lib.MAX_PARTIALS + 1, None)
    with pytest.raises(_lib.HipError, match="op must be"):
        _lib.call("dds_stream", 5, 8, 1, None, 0.0, 1, 8, None)
    with pytest.raises(_lib.HipError, match="no column indices"):
        _lib.call("dds_sellmv", _lib.F64, 64, 1, 1, None, None, 1, 1, 1, 1, None, None, None, None)
    with pytest.raises(_lib.HipError, match="waves must be"):
        _lib.call("dds_mfma_apply", _lib.F16, 4, 1, 1, 1, 1, None, 1, 1, 1, 1, 0, 3, None, None)
    with pytest.raises(_lib.HipError, match="waves must be 4 or 8"):
        _lib.call("dds_mfma_apply_sub", _lib.F16, 4, 1, 1, 1, 1, None, 1, 1, 1, 1, 2, 4, None, None)
    with pytest.raises(_lib.HipError, match="rpw must be"):
        _lib.call("dds_mfma_apply_sub", _lib.F16, 4, 1, 1, 1, 1, None, 1, 1, 1, 1, 8, 3, None, None)
    with pytest.raises(_lib.HipError, match="F16 or BF16"):
        _lib.call("dds_mfma_apply_sub", _lib.F64, 4, 1, 1, 1, 1, None, 1, 1, 1, 1, 8, 4, None, None)
    with pytest.raises(_lib.HipError, match="max_k 1025 exceeds 1024"):
        _lib.call("dds_sym_mfma_apply16", _lib.F16, 4, 1025, 1, 1, 1, 1, None, 1, 1, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="F16 or BF16"):
        _lib.call("dds_sym_mfma_apply16", _lib.F64, 4, 64, 1, 1, 1, 1, None, 1, 1, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="F16 or BF16"):
        _lib.call("dds_sym_mfma_pack16", _lib.F32, 4, 1, 1, 1, None, 1, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="F16 or BF16"):
        _lib.call("dds_mfma_apply", _lib.F64, 4, 1, 1, 1, 1, None, 1, 1, 1, 1, 0, 4, None, None)
    with pytest.raises(_lib.HipError, match="F16 or BF16"):
        _lib.call("dds_mfma_pack_x", _lib.F32, 4, 1, None, 1, 1, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="bad count"):
        _lib.call("dds_pcg_finalize_m", 0, 1, _lib.MAX_PARTIALS + 1, 1, 1, 1, 1, None)
    with pytest.raises(_lib.HipError, match="lanes must be"):
        _lib.call("dds_restrict", _lib.F64, 4, 1, 1, 1, 4, None, None)
    with pytest.raises(_lib.HipError, match="bad grid size"):
        _lib.call("dds_q1_element_grad", 1, 1, 1, 1, None, 1, None)


def test_cpu_tensors_rejected():
    A = torch.eye(3, dtype=torch.float64).to_sparse_csr()
    s = dds.CG(bsr_matmul=False)
    s.setup(A)
    with pytest.raises(ValueError, match="HIP device"):
        s.solve(torch.ones(3, dtype=torch.float64))
    with pytest.raises(ValueError, match="HIP device"):
        dds.gemv_strided_batched(torch.zeros(1, 2, 2, dtype=torch.float64).transpose(1, 2),
                                 torch.zeros(1, 2, dtype=torch.float64))


# ---- setup utilities (known answers of src/dd_solvers/tests/test_bsr.py / test_cusparse.py)

def test_csr_setup_helpers_reject_host_tensors():
    """csr_to_bsr, sort_csr, gather_csr and permute_csr_tensor run on the device, as the
    reference's cuSPARSE / numba kernels do: host tensors are refused, not computed on the
    CPU. Their known answers run on the GPU (tests/test_gpu_setup_helpers.py)."""
    M = torch.sparse_csr_tensor(torch.tensor([0, 2, 3, 5, 9]), torch.tensor([1, 0, 1, 0, 2, 2, 3, 1, 0]),
                                torch.tensor([2, 1, 3, 4, 5, 8, 9, 7, 6], dtype=torch.float64))
    for fn in (lambda: csr_to_bsr(M, 2), lambda: sort_csr(M), lambda: gather_csr(M),
               lambda: permute_csr_tensor(M, torch.arange(4))):
        with pytest.raises(ValueError, match="HIP device"):
            fn()


def test_asm_permutation_dg24():
    A, b, f2s, s2c, k = dg24()
    f2d = torch.arange(24).reshape(8, 3)
    perm = asm_permutation(torch.as_tensor(f2s), torch.as_tensor(s2c), f2d)
    # restated: elements grouped by solver (stable), solvers grouped by coarse cell (stable)
    n_s = s2c.shape[0]
    sd = f2d.numpy()[np.argsort(f2s, kind="stable")].reshape(n_s, -1)
    ref = sd[np.argsort(s2c, kind="stable")].reshape(-1)
    assert perm.tolist() == ref.tolist()
    inv = inverse_permutation(perm)
    assert inv[perm].tolist() == list(range(24))


# ---- problem generators ------------------------------------------------------------------

def test_q1_stencil_closed_form():
    K = P.q1_reference_stiffness()
    for a in range(8):
        for b in range(8):
            dist = bin(a ^ b).count("1")
            assert np.isclose(K[a, b], P.Q1_REF_STIFFNESS_BY_DISTANCE[dist])
    N = 6
    A = P.assemble_q1_3d(N)
    m = N - 1
    c = (2 * m + 2) * m + 2  # an interior node away from the boundary
    row = A.getrow(c).toarray().ravel()
    h = 1.0 / N
    assert np.isclose(row[c], 8 * h / 3)
    assert np.isclose(row[c + 1], 0.0)                  # face neighbour
    assert np.isclose(row[c + 1 + m], -h / 6)           # edge neighbour
    assert np.isclose(row[c + 1 + m + m * m], -h / 12)  # corner neighbour
    assert abs(row.sum()) < 1e-14


def test_q1_structured_matches_element_assembly():
    N = 5
    kappa = P.lognormal_coefficients(N ** 3, 9).reshape(N, N, N)
    A = P.assemble_q1_3d(N, kappa)
    Kref = P.q1_reference_stiffness()
    dofs = P.q1_element_nodes(N)
    n = (N - 1) ** 3
    B = np.zeros((n, n))
    for e in range(N ** 3):
        Ke = kappa.reshape(-1)[e] * Kref / N
        for a in range(8):
            for bb in range(8):
                if dofs[e, a] >= 0 and dofs[e, bb] >= 0:
                    B[dofs[e, a], dofs[e, bb]] += Ke[a, bb]
    assert np.allclose(A.toarray(), B, atol=1e-14)


def test_p1_5_point_stencil():
    A, *_ = P.assemble_p1_2d(8)
    m = 7
    c = 3 * m + 3
    row = A.getrow(c).toarray().ravel()
    assert np.isclose(row[c], 4.0)
    for d in (1, -1, m, -m):
        assert np.isclose(row[c + d], -1.0)
    assert abs(row[c + m + 1]) < 1e-14 and abs(row[c - m - 1]) < 1e-14
    assert np.isclose(row.sum(), 0.0)


def test_sizes_of_the_configurations():
    pb = P.make_problem("p1_2d", 128, box=8, overlap=2)
    d = pb.decomposition
    assert pb.A.shape[0] == 16129 and pb.A.nnz == 111889
    assert d.n_subdomains == 256 and d.sizes.min() == 81 and d.sizes.max() == 144
    d3 = P.box_decomposition(63, 3, 8, 1)
    assert d3.n_subdomains == 512 and d3.sizes.min() == 512 and d3.sizes.max() == 1000
    assert int(d3.sizes.sum()) == 456533


@pytest.mark.parametrize("overlap", [0, 1, 2])
def test_decomposition_invariants(overlap):
    pb = P.make_problem("q1_3d", 10, box=3, overlap=overlap, seed=1, lognormal=True)
    d = pb.decomposition
    n = pb.A.shape[0]
    assert np.array_equal(np.sort(d.perm), np.arange(n))
    seen = np.zeros(n, dtype=int)
    for i in range(d.n_subdomains):
        s = d.subdomain(i)
        assert np.all(np.diff(s) > 0)
        core = np.arange(d.owner_ptr[i], d.owner_ptr[i + 1])
        assert np.isin(core, s).all()
        seen[s] += 1
    assert seen.min() >= 1
    if overlap == 0:
        assert seen.max() == 1
    # the permuted matrix is the original one renumbered
    A0 = pb.A0.toarray()
    assert np.allclose(pb.A.toarray(), A0[d.perm][:, d.perm])


def test_kmeans_decomposition_uneven():
    A, xy, tris, n2d = P.assemble_p1_2d(64, jitter=0.2, seed=3)
    d = P.kmeans_decomposition(A, xy, tris, n2d, 16, overlap=2, seed=3)
    assert d.n_subdomains == 16
    assert d.sizes.max() > d.sizes.min()
    seen = np.zeros(A.shape[0], dtype=int)
    for i in range(d.n_subdomains):
        seen[d.subdomain(i)] += 1
    assert seen.min() >= 1


# ---- symmetric tile layout of the local inverses (host-side model of local_sym.hip) ------

def _unpack_tile(packed, k, V, I, J):
    """Tile (I, J) of the packed layout as a dense (R, 64) block, following the offsets
    documented in csrc/local_sym.hip (diagonal tiles: lower triangle, diagonal halved)."""
    T = -(-k // 64)
    rL = k - 64 * (T - 1)
    R = 64 if I < T - 1 else rL
    nch = 64 // V if J < T - 1 else -(-rL // V)
    D64 = 2048 + 32 * V
    if I < T - 1:
        off = 4096 * (I * (I - 1) // 2) + I * D64 + 4096 * J
    else:
        off = 4096 * ((T - 1) * (T - 2) // 2) + (T - 1) * D64 + 64 * rL * J
    out = np.zeros((R, 64))
    for j in range(nch):
        for r in range(R):
            if J < I:
                o = off + (j * R + r) * V
            elif r >= j * V:
                o = off + V * (j * R - V * j * (j - 1) // 2) + (r - j * V) * V
            else:
                continue
            out[r, j * V:(j + 1) * V] = packed[o:o + V]
    return out


@pytest.mark.parametrize("k", [1, 5, 64, 65, 130, 200])
@pytest.mark.parametrize("V", [2, 4, 8])
def test_sym_layout_and_item_plan(k, V):
    from dd_solvers.solvers import _sym_index, _sym_plan

    rng = np.random.default_rng(k * 10 + V)
    B = rng.standard_normal((k, k))
    A = B + B.T
    idx = _sym_index(k, V)
    flat = A.reshape(-1)
    packed = np.where(idx >= 0, flat[np.maximum(idx, 0)],
                      np.where(idx <= -2, 0.5 * flat[np.maximum(-2 - idx, 0)], 0.0))
    T = -(-k // 64)
    rL = k - 64 * (T - 1)
    D64 = 2048 + 32 * V
    nchL = -(-rL // V)
    diag_last = V * (nchL * rL - V * nchL * (nchL - 1) // 2)
    expect = 4096 * ((T - 1) * (T - 2) // 2) + (T - 1) * D64 + 64 * rL * (T - 1) + diag_last
    assert idx.shape[0] == expect
    # every stored tile matches the matrix: off-diagonal tiles whole, diagonal tiles as their
    # lower triangle with the diagonal halved
    for I in range(T):
        for J in range(I + 1):
            t = _unpack_tile(packed, k, V, I, J)
            blk = A[64 * I:64 * I + t.shape[0], 64 * J:64 * J + 64]
            if J == I:
                blk = np.tril(blk, -1) + np.diag(np.diag(blk)) / 2
            assert np.array_equal(t[:, :blk.shape[1]], blk)
            assert not t[:, blk.shape[1]:].any()
    # work items (pairs of tile rows) + the structured combine reproduce A x
    sizes = np.array([k])
    items, ypos = _sym_plan(sizes)
    x = rng.standard_normal(k)
    xp = np.zeros(64 * T)
    xp[:k] = x
    Yp = np.full(int(ypos[-1]), np.nan)
    for s_, p in items:
        I1, I2 = p, T - 1 - p
        y = np.full(64 * T, np.nan)
        rows = {I1: np.zeros(64), I2: np.zeros(64)}
        hold = np.zeros(64)
        for J in range(I2 + 1):
            contrib = np.zeros(64)
            for I in sorted({I1, I2}, reverse=True):
                if J > I:
                    continue
                t = np.zeros((64, 64))
                tt = _unpack_tile(packed, k, V, I, J)
                t[:tt.shape[0]] = tt
                rows[I] += t @ xp[64 * J:64 * J + 64]
                if J < I:
                    contrib += t.T @ xp[64 * I:64 * I + 64]
                else:
                    rows[I] += t.T @ xp[64 * I:64 * I + 64]
            if J < I2:
                if I1 != I2 and J == I1:
                    hold = contrib
                else:
                    y[64 * J:64 * J + 64] = contrib
        if I1 != I2:
            y[64 * I1:64 * I1 + 64] = rows[I1] + hold
        y[64 * I2:64 * I2 + 64] = rows[I2]
        Yp[ypos[s_] + p * 64 * T: ypos[s_] + (p + 1) * 64 * T] = y
    # combine (sym_combine_kernel): row a sums items p = 0 .. min(P-1, T-1-a//64)
    P = -(-T // 2)
    z = np.array([sum(Yp[ypos[0] + p * 64 * T + a] for p in range(min(P - 1, T - 1 - a // 64) + 1))
                  for a in range(k)])
    assert not np.isnan(z).any()
    assert np.allclose(z, A @ x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 8, 20, 33, 64, 65, 100, 130, 200, 1000])
def test_symm_layout_model(k):
    """The matrix-core symmetric layout (_symm_index) read the way sym_mfma_apply_kernel reads
    it: tile offsets and block indices from the kernel's closed forms, the 16x16x32 MFMA lane
    rules (A: row l%16, k = 8(l/16)+e; B: column l%16, k = 8(l/16)+e; D: rows 4(l/16)+i,
    column l%16), the tile's blocks as the loads into LDS leave them (chunks in swizzled
    slots) and the ds_read_tr16_b64 lane rule (in each group of 16 lanes, lane (r, q) names 4
    elements of a row; lane c receives element c%4 of lane (i, c//4) for i < 4). Items and
    the combine then reproduce A x exactly; every transposed read hits written LDS."""
    from dd_solvers.solvers import _sym_plan, _symm_index

    rng = np.random.default_rng(k)
    B = rng.standard_normal((k, k))
    A = B + B.T
    idx = _symm_index(k)
    flat = A.reshape(-1)
    packed = np.where(idx >= 0, flat[np.maximum(idx, 0)],
                      np.where(idx <= -2, 0.5 * flat[np.maximum(-2 - idx, 0)], 0.0))
    T = -(-k // 64)
    rL = k - 64 * (T - 1)
    nbL = -(-rL // 16)
    last = 512 * (8 * ((T - 1) * (T - 2) // 2) + 6 * (T - 1) + 2 * nbL * (T - 1))
    nkL = -(-rL // 32)
    n_diag_last = sum(1 for Rb in range(nbL) for K2 in range(nkL) if not (K2 == 1 and Rb < 2))
    assert idx.shape[0] == last + 512 * n_diag_last

    def tile_off(I, J):
        if I < T - 1:
            return 512 * (8 * (I * (I - 1) // 2) + 6 * I + 8 * J)
        return 512 * (8 * ((T - 1) * (T - 2) // 2) + 6 * (T - 1) + 2 * nbL * J)

    def blocks(I, J):  # the kernel's issue(): 8 operand registers per lane, absent ones zero
        diag = J == I
        nb = 4 if I < T - 1 else nbL
        nk = -(-rL // 32) if diag and I == T - 1 else 2
        c = np.zeros((8, 64, 8))
        for q in range(8):
            Rb, K2 = q >> 1, q & 1
            if Rb < nb and K2 < nk and not (diag and K2 == 1 and Rb < 2):
                bi = Rb + (max(Rb - 2, 0) if nk == 2 else 0) + K2 if diag else q
                o = tile_off(I, J) + bi * 512
                c[q] = packed[o:o + 512].reshape(64, 8)
        return c

    lanes = np.arange(64)
    xp = np.zeros(64 * T)
    x = rng.standard_normal(k)
    xp[:k] = x

    def row_product(c, xJ):
        D = np.zeros((4, 16, 16))
        for Rb in range(4):
            for K2 in range(2):
                Am = np.zeros((16, 32))
                Bm = np.zeros((32, 16))
                for l in lanes:
                    Am[l % 16, 8 * (l // 16):8 * (l // 16) + 8] = c[2 * Rb + K2][l]
                    if l % 16 == 0:
                        Bm[8 * (l // 16):8 * (l // 16) + 8, 0] = xJ[32 * K2 + 8 * (l // 16):32 * K2 + 8 * (l // 16) + 8]
                D[Rb] += Am @ Bm
        # D lane rule: rows 4(l/16)+i of column 0 sit in lane 16 g, i.e. D[Rb][:, 0]
        return np.concatenate([D[Rb][:, 0] for Rb in range(4)])

    def slot(c):  # symm_slot: LDS slot of chunk c in a block (its own inverse)
        return c ^ (((c >> 4) & 1) << 2)

    def transposed_product(c, xI):
        # the tile as the loads into LDS leave it: block q in 1 KiB, slot s holding chunk
        # slot(s) (16 bytes = 8 elements)
        lds = np.full((8, 64, 8), np.nan)
        for q in range(8):
            for s_ in lanes:
                lds[q, s_] = c[q][slot(s_)]
        flat = lds.reshape(-1)  # element (2-byte) addresses
        g, r, q4 = lanes // 16, (lanes >> 2) & 3, lanes & 3
        out = np.zeros(64)
        for nb in range(4):
            D = np.zeros((16, 16))
            for kr2 in range(2):
                def tr(h):
                    rho = 32 * kr2 + 8 * g + r + 4 * h
                    blk = 2 * (rho >> 4) + (nb >> 1)
                    ln = (rho & 15) + 16 * (2 * (nb & 1) + (q4 >> 1))
                    addr = blk * 512 + np.array([slot(int(v)) for v in ln]) * 8 + 4 * (q4 & 1)
                    got = np.zeros((64, 4))
                    for l in lanes:  # lane c of a 16-lane group: element c % 4 of lane (i, c // 4)
                        gg, cc = l // 16, l % 16
                        for i in range(4):
                            got[l, i] = flat[addr[16 * gg + 4 * i + cc // 4] + cc % 4]
                    return got
                bt = np.concatenate([tr(0), tr(1)], axis=1)
                Am = np.zeros((16, 32))
                Bm = np.zeros((32, 16))
                for l in lanes:
                    if l % 16 == 0:
                        Am[0, 8 * (l // 16):8 * (l // 16) + 8] = xI[32 * kr2 + 8 * (l // 16):32 * kr2 + 8 * (l // 16) + 8]
                    Bm[8 * (l // 16):8 * (l // 16) + 8, l % 16] = bt[l]
                D += Am @ Bm
            out[16 * nb:16 * nb + 16] = D[0]
        return out

    P = -(-T // 2)
    for G in sorted({1, 2, P}):
        # items of G pairs: the sums accumulate per item (yv), each touched row written once
        items, ypos = _sym_plan(np.array([k]), G)
        Q = -(-P // G)
        assert items.shape[0] == Q
        Yp = np.full(int(ypos[-1]), np.nan)
        for s_, q in items:
            I2max = T - 1 - q * G
            yv = np.zeros(64 * (I2max + 1))
            for p in range(q * G, min(P, q * G + G)):
                I1, I2 = p, T - 1 - p
                two = I1 != I2
                rows = {I1: np.zeros(64), I2: np.zeros(64)}
                for J in range(I2 + 1):
                    dt = np.zeros(64)
                    for w, I in ((0, I2), (1, I1)):
                        if w == 1 and not (two and J <= I1):
                            continue
                        c = blocks(I, J)
                        rows[I] += row_product(c, xp[64 * J:64 * J + 64])
                        dt += transposed_product(c, xp[64 * I:64 * I + 64])
                    yv[64 * J:64 * J + 64] += dt
                if two:
                    yv[64 * I1:64 * I1 + 64] += rows[I1]
                yv[64 * I2:64 * I2 + 64] += rows[I2]
            n_rows = min(k, 64 * (I2max + 1))
            o = ypos[s_] + q * 64 * T
            Yp[o:o + n_rows] = yv[:n_rows]
        # the combine (sym_combine_kernel): row a sums items q = 0 .. min(Q-1, (T-1-a//64)//G);
        # with one item (G >= P) that item's vector is the result
        z = np.array([sum(Yp[ypos[0] + q * 64 * T + a] for q in range(min(Q - 1, (T - 1 - a // 64) // G) + 1))
                      for a in range(k)])
        assert not np.isnan(z).any()
        assert np.allclose(z, A @ x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [3, 16, 17, 33, 65, 100, 130])
def test_mfma_operand_layout_model(k):
    """The A-operand order written at setup (_mfma_index) and the B-operand order of the X
    pack, combined with the 16x16x32 MFMA lane rule (A: row l%16, k = 8(l/16)+e; B: column
    l%16, k = 8(l/16)+e; D: rows 4(l/16)+i, column l%16), reproduce A @ X: the operands'
    slot e of lane l holds the same column of A and row of X, so the instruction's own
    numbering of k only permutes the sum."""
    from dd_solvers.solvers import _mfma_index
    rng = np.random.default_rng(k)
    A = rng.standard_normal((k, k))
    X = rng.standard_normal((k, 16))
    idx = _mfma_index(k)
    packed = np.where(idx >= 0, A.reshape(-1)[np.maximum(idx, 0)], 0.0)
    RB, KP = -(-k // 16), -(-k // 32)
    assert packed.shape[0] == RB * KP * 64 * 8
    Bp = np.zeros(KP * 64 * 8)
    for a in range(32 * KP):  # the pack: element (a, c) of X to (K2, lane, slot)
        for c in range(16):
            K2, t2, r16 = a // 32, (a % 32) // 16, a % 16
            lane, i = (r16 // 4) * 16 + c, r16 % 4
            Bp[(K2 * 64 + lane) * 8 + t2 * 4 + i] = X[a, c] if a < k else 0.0
    Y = np.zeros((16 * RB, 16))
    for R in range(RB):
        D = np.zeros((16, 16))
        for K2 in range(KP):
            a0 = (R * KP + K2) * 512
            a_op = packed[a0:a0 + 512].reshape(64, 8)
            b_op = Bp[K2 * 512:K2 * 512 + 512].reshape(64, 8)
            Am = np.zeros((16, 32))
            Bm = np.zeros((32, 16))
            for lane in range(64):
                for e in range(8):
                    Am[lane % 16, 8 * (lane // 16) + e] = a_op[lane, e]
                    Bm[8 * (lane // 16) + e, lane % 16] = b_op[lane, e]
            D += Am @ Bm
        Y[16 * R:16 * R + 16] = D
    assert np.allclose(Y[:k], A @ X, rtol=1e-12, atol=1e-12)


def test_mfma_autotuner_candidates():
    from dd_solvers.tuning import mfma_candidates
    c = mfma_candidates(10_000, 256)
    assert (0, 1, 4) in c and (0, 8, 4) in c and (-1, 8, 8) in c
    assert all(w in (4, 8) and r in (2, 4, 8) for g, w, r in c if g < 0)
    assert all(w in (1, 2, 4, 8) and r == 4 for g, w, r in c if g >= 0)
    assert all(g <= 0 or g * w < 10_000 for g, w, _ in c)
    small = mfma_candidates(10, 256)
    assert [x for x in small if x[0] >= 0] == [(0, 1, 4), (0, 2, 4), (0, 4, 4), (0, 8, 4)]
    assert len([x for x in small if x[0] < 0]) == 6


def test_sell_compact_column_format():
    """The compact SELL columns decode to the int32 columns: 16-bit offsets from a per-slot
    base where a slot spans < 65536 columns, int32 for the slices holding a wider slot."""
    import torch
    from dd_solvers.solvers import _SELLOperator
    rng = np.random.default_rng(0)
    width = torch.tensor([3, 1, 4, 2, 5])
    n_slots = int(width.sum())
    col = rng.integers(0, 60000, size=(n_slots, 64))
    col[4] += 200_000          # slot 4 (slice 2) lies far away but narrow: still compact
    col[5, 7] = 900_000        # slot 5 (slice 2) spans > 65535: slice 2 goes wide
    col[10, :] = 70_000 + np.arange(64)
    op = object.__new__(_SELLOperator)
    op.col = torch.as_tensor(col.reshape(-1), dtype=torch.int32)
    op.device = torch.device("cpu")
    op._compact_columns(width)
    cp = op.col_ptr.numpy()
    assert (cp < 0).tolist() == [False, False, True, False, False]
    d16 = op.col16.numpy().view(np.uint16).astype(np.int64)
    c32 = op.col32.numpy()
    base = op.slot_base.numpy()
    slot0 = np.concatenate([[0], np.cumsum(width.numpy())])
    for q in range(5):
        for j in range(int(width[q])):
            slot = slot0[q] + j
            got = (base[slot] + d16[cp[q] + 64 * j + np.arange(64)]) if cp[q] >= 0 else \
                c32[~cp[q] + 64 * j + np.arange(64)]
            assert np.array_equal(got, col[slot]), (q, j)
    assert d16.shape[0] == 64 * (n_slots - 4) and c32.shape[0] == 64 * 4


def test_sell_slot_pair_order():
    """_pair_slots moves every stored array into slot pairs; read back with the kernels'
    addressing (slot pair p of lane l at 128 p + 2 l (+1), an odd last slot j at 64 j + l;
    a lane's two 16-bit offsets as one little-endian word) it gives each slot's values and
    columns, for even and odd widths and for compact and wide slices."""
    import torch
    from dd_solvers.solvers import _SELLOperator
    rng = np.random.default_rng(1)
    width = torch.tensor([3, 1, 4, 2, 5, 6])
    n_slots = int(width.sum())
    col = rng.integers(0, 60000, size=(n_slots, 64))
    col[5, 7] = 900_000        # slice 2 goes wide (int32 columns)
    val = rng.standard_normal((n_slots, 64))
    op = object.__new__(_SELLOperator)
    op.col = torch.as_tensor(col.reshape(-1), dtype=torch.int32)
    op.values = torch.as_tensor(val.reshape(-1))
    op.device = torch.device("cpu")
    op.slice_ptr = torch.as_tensor(np.concatenate([[0], np.cumsum(64 * width.numpy())]))
    op.stored = 64 * n_slots
    op._compact_columns(width)
    op._pair_slots(width)
    cp = op.col_ptr.numpy()
    words = op.col16.numpy().view(np.uint32) if op.col16.numel() else None
    d16 = op.col16.numpy().view(np.uint16).astype(np.int64)
    c32, base = op.col32.numpy(), op.slot_base.numpy()
    v, call = op.values.numpy(), op.col.numpy()
    slot0 = np.concatenate([[0], np.cumsum(width.numpy())])
    lanes = np.arange(64)
    for q in range(width.shape[0]):
        w, s0 = int(width[q]), int(op.slice_ptr[q])
        for j in range(w):
            slot = slot0[q] + j
            paired = j < 2 * (w // 2)
            pos = 128 * (j // 2) + 2 * lanes + j % 2 if paired else 64 * j + lanes
            assert np.array_equal(v[s0 + pos], val[slot]), (q, j)
            assert np.array_equal(call[s0 + pos], col[slot]), (q, j)
            if cp[q] >= 0:
                if paired:
                    wd = words[(cp[q] + 128 * (j // 2)) // 2 + lanes].astype(np.int64)
                    off = (wd >> 16) if j % 2 else (wd & 0xFFFF)
                else:
                    off = d16[cp[q] + pos]
                got = base[slot] + off
            else:
                got = c32[~cp[q] + pos]
            assert np.array_equal(got, col[slot]), (q, j)


def _symm_schedule(k, G, q, NS, NT):
    """Restates the control flow of sym_mfma_apply_kernel for one item (subdomain of k
    unknowns, pairs q G .. q G + G - 1, tile rows p and T-1-p of each in increasing order)
    and checks, with vector memory operations completing in issue order: the scales (and the
    item's terms) are in before the loop; each wait count inside the loop (LPT (NS - 1), LPT = 8 + 2 XL loads per
    tile: XL = ceil(NT 128 / 256) with the x terms per tile (NT = 3), else 0 with the item's
    terms loaded first) equals the loads issued after the awaited tile's; the awaited tile
    is still in its stage and is the tile computed, with the x terms of its block J and, at
    the first tile of a row, of its block I; a stage is reloaded only after
    its task was computed; the item reads every tile of its rows once, a whole subdomain in
    storage order."""
    T = -(-k // 64)
    P = -(-T // 2)
    p0, p_end = q * G, min(P, q * G + G)
    I2max = T - 1 - p0
    nlow = p_end - p0
    nrows = nlow + (T - p0) - max(T - p_end, p_end)

    def row(i):
        return p0 + i if i < nlow else max(T - p_end, p_end) + (i - nlow)

    rows = [row(i) for i in range(nrows)]
    assert sorted(rows) == sorted({p for p in range(p0, p_end)} | {T - 1 - p for p in range(p0, p_end)})
    assert rows == sorted(rows)

    def advance(c):
        i, I, J = c
        J += 1
        if J > I:
            i, J = i + 1, 0
            I = row(i)
        return (i, I, J)

    NF = NS - 1
    XL = -(-NT * 128 // 256) if NT > 2 else 0  # terms per tile, or the item's before its tiles
    LPT = 8 + 2 * XL
    ntot = sum(I + 1 for I in rows)
    xdw = 0 if XL else (I2max + 1) * NT * 128 // 4
    queue = [("sig",)] * ((xdw + 63) // 64 + (I2max + 1 + 63) // 64)  # terms, then scales
    stage_task = {}

    def issue(t, c):
        prev = stage_task.get(t % NS)
        real = c[0] < nrows
        I, J = (c[1], c[2]) if real else (None, None)
        queue.extend([("tile", t, (I, J) if real else None)] * 8)  # 8 block loads, always
        queue.extend([("xJ", t, J)] * XL)
        queue.extend([("xI", t, I if real and J == 0 else None)] * XL)
        stage_task[t % NS] = t
        return prev, (advance(c) if real else c)

    ic = (0, row(0), 0)
    for t in range(NF):
        _, ic = issue(t, ic)
    assert max(i for i, e in enumerate(queue) if e == ("sig",)) <= len(queue) - 1 - LPT * NF
    cc = (0, row(0), 0)
    seen = []
    for t in range(ntot):
        prev, ic = issue(t + NF, ic)
        assert prev is None or prev < t             # the stage's task was computed
        mine = [i for i, e in enumerate(queue) if len(e) > 1 and e[1] == t]
        assert len(mine) == LPT
        assert len(queue) - 1 - max(mine) == LPT * NF  # the exact count after t's loads
        assert stage_task[t % NS] == t
        got = {queue[i][0]: queue[i][2] for i in mine}
        assert got["tile"] == (cc[1], cc[2])        # the tile loaded is the task computed
        if XL:
            assert got["xJ"] == cc[2]
            if cc[2] == 0:                          # a row starts: its x_I operands
                assert got["xI"] == cc[1]
        assert cc[2] <= cc[1] <= I2max
        seen.append((cc[1], cc[2]))
        cc = advance(cc)
    assert cc[0] == nrows
    assert seen == [(I, J) for I in rows for J in range(I + 1)]
    if G >= P:
        assert seen == [(I, J) for I in range(T) for J in range(I + 1)]  # storage order


@pytest.mark.parametrize("NS,NT", [(3, 2), (3, 3), (2, 2), (2, 3), (4, 3)])
def test_symm_schedule(NS, NT):
    for k in (1, 63, 64, 65, 127, 128, 129, 700, 1000, 2049, 4096):
        T = -(-k // 64)
        P = -(-T // 2)
        for G in sorted({1, 2, 3, P}):
            for q in range(-(-P // G)):
                _symm_schedule(k, G, q, NS, NT)


def _sym16_schedule(k, NF, NT):
    """Restates the control flow of sym16_apply_kernel for one subdomain of k unknowns (4
    wavefronts, NS = NF + 2 LDS stages, one barrier per tile) and checks, with each
    wavefront's vector memory operations completing in issue order: every wavefront issues
    the same loads per tile (2 tile blocks, NT of the 4 NT packed-X chunks) and its wait
    count (LPT NF) leaves exactly the loads of the tiles after the awaited one in flight; the
    scales are in before the first tile; the four quarters of a stage are the tile computed,
    its X_J and, at the first tile of a row, its X_I; a stage is refilled only after the
    barrier that follows its last read; the walk (odd rows right to left) reads every tile
    (I, J <= I) once and starts and ends every row once."""
    T = -(-k // 64)
    NS = NF + 2
    LPT = 2 + NT
    ntot = T * (T + 1) // 2

    def first(I, J):
        return J == (I if I & 1 else 0)

    def last(I, J):
        return J == (0 if I & 1 else I)

    def advance(I, J):
        if I & 1:
            J -= 1
            if J < 0:
                I, J = I + 1, 0
        else:
            J += 1
            if J > I:
                I, J = I + 1, I + 1
        return I, J

    queues = [[("sig",)] for _ in range(4)]  # per wavefront, in issue order
    stage = {}          # stage -> (tile index, (I, J)) it holds
    last_read = {}      # stage -> tile index computed from it
    barrier = -1        # the last barrier passed: all computes before it are done

    def issue(t, c):
        real = t < ntot
        st = t % NS
        if st in last_read:  # refilled only after the barrier following its last read
            assert last_read[st] < barrier, (t, st, last_read[st], barrier)
        for w in range(4):
            q = queues[w]
            q += [("tile", t, (c, 2 * w + u) if real else None) for u in range(2)]
            for u in range(NT):
                ch = w * NT + u
                if ch < 2 * NT:
                    q.append(("x", t, ("J", c[1], ch) if real else None))
                else:
                    q.append(("x", t, ("I", c[0], ch - 2 * NT) if real and first(*c) else None))
        stage[st] = (t, c if real else None)
        return advance(*c) if real else c

    ic = (0, 0)
    for t in range(NF):
        ic = issue(t, ic)
    cc = (0, 0)
    seen, firsts, lasts = [], [], []
    for t in range(ntot):
        ic = issue(t + NF, ic)
        got = {}
        for w in range(4):
            q = queues[w]
            mine = [i for i, e in enumerate(q) if len(e) > 1 and e[1] == t]
            assert len(mine) == LPT
            assert len(q) - 1 - max(mine) == LPT * NF  # the wait leaves only later tiles' loads
            if t == 0:
                assert max(i for i, e in enumerate(q) if e == ("sig",)) < min(mine)
            for i in mine:
                got.setdefault(q[i][0], []).append(q[i][2])
        barrier = t  # every wavefront's loads of t are in, every compute of t - 1 is done
        assert stage[t % NS] == (t, cc)
        assert sorted(b for _, b in got["tile"]) == list(range(8))
        assert all(c == cc for c, _ in got["tile"])
        xs = [x for x in got["x"] if x is not None]
        assert sorted(ch for kind, _, ch in xs if kind == "J") == list(range(2 * NT))
        assert all(blk == cc[1] for kind, blk, _ in xs if kind == "J")
        xi = sorted(ch for kind, blk, ch in xs if kind == "I" and blk == cc[0])
        assert xi == (list(range(2 * NT)) if first(*cc) else [])
        last_read[t % NS] = t
        seen.append(cc)
        if first(*cc):
            firsts.append(cc[0])
        if last(*cc):
            lasts.append(cc[0])
        cc = advance(*cc)
    assert sorted(seen) == sorted((I, J) for I in range(T) for J in range(I + 1))
    assert firsts == list(range(T)) and lasts == list(range(T))
    assert [I for I, _ in seen] == sorted(I for I, _ in seen)  # rows in increasing order


@pytest.mark.parametrize("NF,NT", [(1, 2), (1, 3), (2, 2), (2, 3)])
def test_sym16_schedule(NF, NT):
    for k in (1, 63, 64, 65, 128, 129, 700, 1000, 1024):
        _sym16_schedule(k, NF, NT)


def test_symm_group_choice():
    from dd_solvers.solvers import _sym_plan, _symm_group, _sym_written_rows
    many = np.full(4096, 1000)
    assert _symm_group(many, 256) == 8          # whole subdomains: one item each
    items, ypos = _sym_plan(many, 8)
    assert items.shape[0] == 4096
    few = np.full(64, 4000)                     # T = 63, P = 32 pairs
    g = _symm_group(few, 256)
    assert g == 1 or 64 * -(-32 // g) >= 32 * 256
    items, ypos = _sym_plan(few, g)
    assert _sym_written_rows(few, g) <= int(ypos[-1])
    # items longest first
    T = 63
    work = [sum((T + 1) if p != T - 1 - p else (T - p) for p in range(q * g, min(32, q * g + g)))
            for _, q in items]
    assert work == sorted(work, reverse=True)


def _pivoted_gj_model(A):
    """The steps of pivoted_inverse_kernel (csrc/setup.hip) on one matrix."""
    a = np.array(A, dtype=np.float64)
    k = a.shape[0]
    pr = np.zeros(k, dtype=np.int64)
    for q in range(k):
        col = np.abs(a[q:, q])
        if not (col.max() > 0):
            return None, q + 1
        p = q + int(np.argmax(col))  # the first of equal maxima
        pr[q] = p
        a[[p, q]] = a[[q, p]]
        ipiv = 1.0 / a[q, q]
        colq = a[:, q].copy()
        a[q] = a[q] * ipiv
        a[q, q] = ipiv
        for i in range(k):
            if i != q:
                rowi = a[i].copy()
                rowi[q] = 0.0
                a[i] = rowi - colq[i] * a[q]
    for q in range(k - 1, -1, -1):
        p = pr[q]
        a[:, [p, q]] = a[:, [q, p]]
    return a, 0


@pytest.mark.parametrize("kind", ["random", "zero_diagonal", "permutation", "indefinite", "singular"])
def test_pivoted_inverse_model(kind):
    """Gauss-Jordan with partial pivoting as the kernel runs it gives the inverse (numpy),
    including matrices that cannot be inverted without row exchanges, and flags a singular
    one at the first column without a pivot."""
    rng = np.random.default_rng(7)
    k = 9
    if kind == "random":
        A = rng.standard_normal((k, k))
    elif kind == "zero_diagonal":
        A = rng.standard_normal((k, k))
        np.fill_diagonal(A, 0.0)
    elif kind == "permutation":
        A = np.eye(k)[rng.permutation(k)] * rng.uniform(1, 2, k)
    elif kind == "indefinite":
        B = rng.standard_normal((k, k))
        A = B + B.T
    else:
        A = rng.standard_normal((k, k))
        A[:, 4] = 0.0
    inv, info = _pivoted_gj_model(A)
    if kind == "singular":
        assert info == 5 and inv is None
    else:
        assert info == 0
        assert np.allclose(inv, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
