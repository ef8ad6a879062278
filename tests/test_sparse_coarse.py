"""Sparse coarse solver (dd_solvers.sparse_coarse) on the host: the library's native
symbolic analysis (dds_nd_plan_build) against its restatement (oracle.nd_plan.NDPlan) table
by table, and the plan with the restated block LDL^T factor and sweeps
(oracle.nd_plan.factorize_host / solve_host, same tables) against scipy's sparse direct
solve (SuperLU) of the same matrix. No GPU is needed."""

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla
import torch

from dd_solvers import _lib, dg
from dd_solvers import sparse_coarse as sc
from oracle import nd_plan as onp

PLAN_FIELDS = ("perm", "fj0", "fnj", "fns", "level_ptr", "foff", "wtoff", "soff", "wbase",
               "moff", "maps", "push_dst", "push_src", "push_ptr", "uoff2", "fac_size",
               "slot_size", "xs_size", "nT", "d_top", "n_fronts")


def same_plan(p, q):
    """Every table of the native plan p equals the restatement's q."""
    for f in PLAN_FIELDS:
        assert np.array_equal(np.asarray(getattr(p, f)), np.asarray(getattr(q, f))), f
    assert len(p.fwd) == len(q.fwd) and len(p.bwd) == len(q.bwd)
    for (x, lx), (y, ly) in zip(p.fwd + p.bwd, q.fwd + q.bwd):
        assert np.array_equal(x, y) and lx == ly
    assert all(np.array_equal(a, b) for a, b in zip(p.S, q.S))
    assert p.children == [[int(c) for c in cs] for cs in q.children]
    assert all((a is None and b is None) or np.array_equal(a, b) for a, b in zip(p.cmap, q.cmap))
    assert np.array_equal(p.Ap.indptr, q.Ap.indptr) and np.array_equal(p.Ap.indices, q.Ap.indices)
    assert np.array_equal(p.Ap.data, q.Ap.data)


def lap2d(nx, ny):
    T = lambda m: sps.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    return (sps.kron(sps.identity(ny), T(nx)) + sps.kron(T(ny), sps.identity(nx))).tocsr()


def dg_coarse(fine, coarse):
    """A_C = R_C A R_C^T of the reference's DG experiment (p = 1, solvers = coarse mesh)."""
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(fine), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, coarse, coarse)
    n_c = spc.shape[0]
    cell = np.repeat(np.arange(n_c), k * spc)
    R = sps.csr_matrix((np.ones(A.shape[0]), (cell, np.arange(A.shape[0]))), shape=(n_c, A.shape[0]))
    return (R @ A @ R.T).tocsr()


def random_spd_pieces(seed=0):
    """Block-diagonal SPD matrix of three disconnected random sparse pieces."""
    rng = np.random.default_rng(seed)
    blocks = []
    for m in (37, 90, 5):
        B = sps.random(m, m, density=0.05, random_state=rng)
        B = B + B.T
        blocks.append(B + sps.diags(np.abs(B).sum(1).A1 + 1.0))
    return sps.block_diag(blocks).tocsr()


def _check(A, leaf, top_max, tol=1e-11):
    plan = sc.NDPlan(A, leaf_size=leaf, top_max=top_max)
    plan.validate()
    same_plan(plan, onp.NDPlan(A, leaf_size=leaf, top_max=top_max))
    fac = onp.factorize_host(plan, plan.Ap.data)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = onp.solve_host(plan, fac, b)
    xr = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xr) <= tol * np.linalg.norm(xr)
    return plan


@pytest.mark.parametrize("A,leaf,top_max", [
    (lap2d(1, 1), 32, 4096),            # one unknown
    (lap2d(7, 5), 4, 8),                # several levels, small top
    (lap2d(40, 33), 16, 300),
    (lap2d(40, 33), 16, 0),             # no top front: the roots are ordinary fronts
    (random_spd_pieces(), 8, 16),       # disconnected graph
])
def test_sweeps_solve_like_superlu(A, leaf, top_max):
    _check(A, leaf, top_max)


def test_dg_coarse_matrix():
    """The coarse matrix of the reference's DG experiment, S6 fine / S4 coarse (512 cells),
    split down to 8-cell leaves under a 64-unknown top front."""
    A = dg_coarse(6, "S4")
    plan = _check(A, 8, 64)
    assert plan.n_levels >= 4 and plan.nT <= 64


def test_plan_structure():
    """Nested dissection: perm is a permutation; every coupling set lies above its front;
    an edge of A never joins two fronts unless one is an ancestor of the other (so the
    factor has no fill outside the fronts)."""
    A = lap2d(30, 21)
    plan = sc.NDPlan(A, leaf_size=8, top_max=40)
    assert np.array_equal(np.sort(plan.perm), np.arange(A.shape[0]))
    owner = np.empty(A.shape[0], dtype=np.int64)
    for t in range(plan.n_fronts):
        a, nj = plan.fj0[t], plan.fnj[t]
        owner[a:a + nj] = t
        assert np.all(plan.S[t] >= a + nj)
    parent = np.full(plan.n_fronts, -1)
    for t in range(plan.n_fronts):
        for c in plan.children[t]:
            parent[c] = t

    def ancestors(t):
        out = set()
        while parent[t] >= 0:
            t = parent[t]
            out.add(t)
        return out

    Ap = plan.Ap.tocoo()
    for i, j in zip(Ap.row, Ap.col):
        ti, tj = owner[i], owner[j]
        if ti != tj:
            assert tj in ancestors(ti) or ti in ancestors(tj), (i, j)
    # heights: children are processed in earlier launches than their parents
    lev = np.searchsorted(plan.level_ptr, np.arange(plan.n_fronts), side="right") - 1
    for t in range(plan.n_fronts):
        for c in plan.children[t]:
            assert lev[c] < lev[t]


def test_planted_mistake_is_caught():
    """A factor entry off by 1e-6 (relative) is visible to the check above."""
    A = lap2d(20, 20)
    plan = sc.NDPlan(A, leaf_size=8, top_max=30)
    fac = onp.factorize_host(plan, plan.Ap.data)
    t = plan.n_fronts // 2
    nj, ns, foff = int(plan.fnj[t]), int(plan.fns[t]), int(plan.foff[t])
    assert ns > 0
    fac[foff + nj] *= 1 + 1e-6               # W[0, 0] (column-major stacked factor)
    b = np.random.default_rng(1).standard_normal(A.shape[0])
    x = onp.solve_host(plan, fac, b)
    xr = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xr) > 1e-11 * np.linalg.norm(xr)


def test_not_positive_definite():
    A = lap2d(10, 10) - 3.0 * sps.identity(100)
    plan = sc.NDPlan(A.tocsr(), leaf_size=8, top_max=20)
    with pytest.raises(np.linalg.LinAlgError):
        onp.factorize_host(plan, plan.Ap.data)


def test_front_capacity_is_enforced():
    with pytest.raises(ValueError, match="exceeds"):
        sc.NDPlan(lap2d(8000, 1), leaf_size=8000, top_max=0)   # one 8000-unknown leaf


def test_host_validation_of_the_sweep_entry_points():
    with pytest.raises(_lib.HipError, match="front of"):
        _lib.call("dds_nd_forward", _lib.F64, _lib.F64, 1, 1, 1, 1, 1, 1, 1, 1, 7681, None, None)
    with pytest.raises(_lib.HipError, match="front of"):
        _lib.call("dds_nd_backward", _lib.F64, _lib.F64, 1, 1, 1, 1, 1, 1, 1, 1, 1, 9000, None, None)
    with pytest.raises(_lib.HipError, match="unsupported"):
        _lib.call("dds_nd_forward", _lib.F16, _lib.F16, 1, 1, 1, 1, 1, 1, 1, 1, 8, None, None)
    with pytest.raises(_lib.HipError, match="unsupported"):  # FP32 vectors, FP64 factor
        _lib.call("dds_nd_backward", _lib.F32, _lib.F64, 1, 1, 1, 1, 1, 1, 1, 1, 1, 8, None, None)
    with pytest.raises(_lib.HipError, match="bad sizes"):
        _lib.call("dds_nd_top_assemble", _lib.F64, 0, -1, 0, 0, 1, 1, 1, 1, None, None)
    with pytest.raises(_lib.HipError, match="null pointer"):
        _lib.call("dds_nd_top_push", _lib.F64, 0, 4, 2, 1, 1, None, None, 1, 1, None, None)


def test_cpu_tensors_rejected():
    with pytest.raises(ValueError, match="HIP device"):
        sc.SparseCholesky().setup(torch.eye(3, dtype=torch.float64).to_sparse_csr())


# ---------------------------------------------------------------------------------------
# stored zeros; the local-solver slot (CSRBatch, block-diagonal plans)
# ---------------------------------------------------------------------------------------

def test_stored_zeros_leave_the_pattern():
    """A 3D Q1 stiffness matrix stores its nearest-neighbour couplings as exact zeros, and
    the dissection ignores stored zeros: the fronts must ignore them as well (else a child's
    coupling set reaches into a sibling subtree). Also a planted zero linking two corners."""
    from dd_solvers import problems as P
    pb = P.make_problem("q1_3d", 10, box=5, overlap=0, seed=0, lognormal=True)
    A = pb.A.tocsr()
    assert np.count_nonzero(A.data == 0) > 0
    _check(A, 16, 0)
    _check(A, 16, 64)
    C = lap2d(20, 20).tocoo()
    L = sps.csr_matrix((np.concatenate([C.data, [0.0, 0.0]]),
                        (np.concatenate([C.row, [0, 399]]), np.concatenate([C.col, [399, 0]]))),
                       shape=C.shape)
    assert L[0, 399] == 0 and 399 in L.indices[L.indptr[0]:L.indptr[1]]
    _check(L, 8, 32)


def _sorted_spd(n, seed):
    rng = np.random.default_rng(seed)
    B = sps.random(n, n, density=0.15, random_state=rng)
    A = (B + B.T + sps.diags(np.full(n, float(n)))).tocsr()
    A.sort_indices()
    return A


def test_local_csr_batch_oracle_pinned():
    """The oracle's restatement of the reference's CSRBatch construction against scipy's
    slicing of the diagonal blocks; a column index moved inside a block is caught."""
    from oracle.kernels import local_csr_batch as ref
    A, k = _sorted_spd(60, 3), 12
    nrows, nnz, vals, cols, ro = ref(A.indptr, A.indices, A.data, k)
    assert np.all(nrows == k) and ro.shape[0] == 60 + 5

    def blocks_match(vals, cols):
        for i in range(5):
            e0 = int(nnz[:i].sum())
            B = sps.csr_matrix((vals[e0:e0 + nnz[i]], cols[e0:e0 + nnz[i]],
                                ro[i * (k + 1):(i + 1) * (k + 1)]), shape=(k, k))
            if (B != A[i * k:(i + 1) * k, i * k:(i + 1) * k]).nnz:
                return False
        return True

    assert blocks_match(vals, cols)
    bad = cols.copy()
    bad[3] = (bad[3] + 1) % k
    assert not blocks_match(vals, bad)


def test_local_csr_batch_matches_reference_layout():
    """The CSRBatch the Schwarz set-up hands a sparse local solver (built from the matrix
    entries in any order) equals the reference's layout field by field."""
    from oracle.kernels import local_csr_batch as ref
    from dd_solvers.solvers import local_csr_batch
    A, k = _sorted_spd(72, 4), 8
    C = A.tocoo()
    r, c, v = C.row.astype(np.int64), C.col.astype(np.int64), C.data
    keep = r // k == c // k
    o = np.random.default_rng(0).permutation(int(keep.sum()))
    s, a, b = (r // k)[keep][o], (r % k)[keep][o], (c % k)[keep][o]
    T = lambda x: torch.as_tensor(x)
    got = local_csr_batch(torch.arange(0, 73, k), T(s), T(a), T(b), T(v[keep][o]), torch.float64)
    want = ref(A.indptr, A.indices, A.data, k)
    for g, w in zip(got, want):
        assert np.array_equal(g.numpy(), w)
    assert got.nrows.dtype == torch.int32 and got.col_indices.dtype == torch.int32
    assert got.row_offsets.dtype == torch.int32 and got.nnz.dtype == torch.int64


def _overlapped_blocks():
    from dd_solvers import problems as P
    pb = P.make_problem("p1_2d", 16, box=4, overlap=2, seed=0)
    d, A = pb.decomposition, pb.A.tocsr()
    subs = [A[idx][:, idx].tocsr() for idx in
            (d.sub_idx[d.sub_ptr[i]:d.sub_ptr[i + 1]] for i in range(d.n_subdomains))]
    return d, subs


def test_local_csr_batch_uneven_overlapping_round_trip():
    """Uneven, overlapping subdomains: the CSRBatch read back as a block-diagonal matrix is
    diag(A[S_i][:, S_i])."""
    from dd_solvers.solvers import local_csr_batch
    d, subs = _overlapped_blocks()
    assert len({m.shape[0] for m in subs}) > 1
    s, a, b, v = [], [], [], []
    for i, m in enumerate(subs):
        C = m.tocoo()
        s.append(np.full(C.nnz, i))
        a.append(C.row)
        b.append(C.col)
        v.append(C.data)
    T = lambda x: torch.as_tensor(np.concatenate(x).astype(np.int64))
    batch = local_csr_batch(torch.as_tensor(d.sub_ptr), T(s), T(a), T(b),
                            torch.as_tensor(np.concatenate(v)), torch.float64)
    H = sc.csr_batch_to_host(batch)
    assert (H != sps.block_diag(subs).tocsr()).nnz == 0
    assert np.array_equal(batch.nrows.numpy(), np.diff(d.sub_ptr))


def test_block_diagonal_plan():
    """The plan of a batch (diag of the subdomain matrices): no top front, every front
    inside one subdomain, and the sweeps solve each block like SuperLU."""
    d, subs = _overlapped_blocks()
    B = sps.block_diag(subs).tocsr()
    plan = _check(B, 16, 0)
    assert not plan.has_top
    block = np.repeat(np.arange(len(subs)), [m.shape[0] for m in subs])
    for t in range(plan.n_fronts):
        own = block[plan.perm[plan.fj0[t]:plan.fj0[t] + plan.fnj[t]]]
        cpl = block[plan.perm[plan.S[t]]]
        assert np.all(own == own[0]) and np.all(cpl == own[0])


def test_mixed_factor_dtype_needs_no_top_front():
    """An FP32 factor under FP64 vectors is for plans without a dense top front (the top
    front's kernels take one dtype); the check comes before any device work."""
    solver = sc.SparseCholesky(leaf_size=8, top_max=64)
    solver.dtype, solver.fac_dtype = torch.float64, torch.float32
    with pytest.raises(ValueError, match="without a top front"):
        solver._setup_host(lap2d(30, 30), "cpu")


def test_plan_cache_same_pattern_new_values():
    """A second matrix with the same pattern reuses the plan, and its own values come back
    in the plan's order: its factor solves it like SuperLU. Another pattern gets a new plan."""
    A1 = lap2d(30, 24)
    A2 = A1.copy()
    rng = np.random.default_rng(3)
    A2.data = A2.data * (1.0 + 0.1 * rng.random(A2.nnz))
    A2 = ((A2 + A2.T) * 0.5).tocsr()
    A2 = A2 + sps.diags(np.abs(A2).sum(1).A1)
    A2 = A2.tocsr()
    p1, v1 = sc.plan_for(A1, 16, 64)
    p2, v2 = sc.plan_for(A2, 16, 64)
    assert p2 is p1 and not np.array_equal(v1, v2)
    fac = onp.factorize_host(p2, v2)
    b = rng.standard_normal(A2.shape[0])
    x = onp.solve_host(p2, fac, b)
    xr = spla.spsolve(A2.tocsc(), b)
    assert np.linalg.norm(x - xr) <= 1e-11 * np.linalg.norm(xr)
    p3, _ = sc.plan_for(lap2d(31, 24), 16, 64)
    assert p3 is not p1
    sc.clear_plan_cache()
    p4, _ = sc.plan_for(lap2d(31, 24), 16, 64)
    assert p4 is not p3


def test_item_shapes_on_wide_fronts():
    """A plan with long coupling sets on levels of few workgroups: forward items split into
    column parts (R < 256) and backward items narrowed to 32 or 16 columns; every item shape
    is one the kernels take, and the host restatement over those items solves like SuperLU."""
    I = sps.identity(24)
    T = sps.diags([-np.ones(23), 2 * np.ones(24), -np.ones(23)], [-1, 0, 1])
    A = (sps.kron(sps.kron(T, I), I) + sps.kron(sps.kron(I, T), I)
         + sps.kron(sps.kron(I, I), T)).tocsr()          # 3D Laplacian: wide separators
    plan = _check(A, 32, 16)
    R = np.concatenate([it[:, 10] for it, _ in plan.fwd])
    C = np.concatenate([it[:, 9] for it, _ in plan.bwd])
    ns_b = np.concatenate([it[:, 2] for it, _ in plan.bwd])
    assert set(np.unique(R)) <= {64, 128, 256} and (R < 256).any()
    assert set(np.unique(C)) <= {16, 32, 64}
    assert (C[ns_b > 256] == 16).all() and (ns_b > 256).any()
    assert (C[(ns_b > 128) & (ns_b <= 256)] <= 32).all()


def test_native_plan_independent_of_threads(monkeypatch):
    """The dissection runs its pieces on a thread pool (OMP_NUM_THREADS threads, at most
    16); the plan is the same for any thread count, and equal to the restatement's."""
    A = sps.block_diag([lap2d(150, 130), random_spd_pieces(4), lap2d(33, 1)]).tocsr()
    ref = onp.NDPlan(A, leaf_size=16, top_max=200)
    for t in ("1", "3", "8"):
        monkeypatch.setenv("OMP_NUM_THREADS", t)
        same_plan(sc.NDPlan(A, leaf_size=16, top_max=200), ref)


def test_native_plan_rejects_bad_patterns():
    A = lap2d(5, 5)
    indptr = A.indptr.astype(np.int64)
    bad = A.indices.astype(np.int32).copy()
    bad[[0, 1]] = bad[[1, 0]]                      # unsorted row
    import ctypes
    h = ctypes.c_void_p()
    st = _lib._lib.dds_nd_plan_build(25, indptr.ctypes.data, bad.ctypes.data, 4, 8, ctypes.byref(h))
    assert st == 1 and b"sorted" in _lib._lib.dds_last_error() and not h.value
    st = _lib._lib.dds_nd_plan_build(25, indptr.ctypes.data, A.indices.astype(np.int32).ctypes.data,
                                     0, 8, ctypes.byref(h))
    assert st == 1 and b"bad arguments" in _lib._lib.dds_last_error()
    assert _lib._lib.dds_nd_plan_field(None, 1, None) == -1
    with pytest.raises(ValueError, match="top_max"):
        sc.SparseCholesky(top_max=sc.TOP_CAP + 1)
