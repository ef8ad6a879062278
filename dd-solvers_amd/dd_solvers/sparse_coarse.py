"""Sparse SPD coarse solver: nested-dissection block LDL^T with level-scheduled HIP sweeps.

The reference's coarse slot is `CUDSS()` (src/dd_solvers/dd_solvers/solvers.py:85-113), a
sparse Cholesky of the coarse matrix A_C. Its experiments use coarse spaces of up to
millions of cells (one per cell of the coarse mesh), where a dense factor is out of the
question. `SparseCholesky` fills that slot without cuDSS:

Ordering (host, native, setup). Nested dissection of the graph of A_C by recursive bisection:
from a pseudo-peripheral node (the last node of a breadth-first search), the level set of a
second breadth-first search that splits the nodes most evenly with the fewest nodes is the
separator; the two sides are dissected in turn down to `leaf_size` nodes, and disconnected
pieces become sibling subtrees. Tree nodes at depth < d (the largest d whose separators
hold at most `top_max` nodes together) form one dense top front T, numbered last; the
subtrees below it are numbered in post-order, so every front's unknowns J are one
contiguous range and its coupling set S (the unknowns of ancestors that J touches after
elimination) lies above it.

Factorisation (device, setup). Multifrontal, one dds_nd_factor launch per height of the
tree, one workgroup per front: the front F of J is assembled in a workspace from the entries
of A_C in the columns J and the children's update matrices (added child by child, so the
sums are deterministic), and a Gauss-Jordan sweep of its first nJ pivots (no pivoting: the
pivots of an SPD matrix are positive) leaves in place

    Finv = F_JJ^-1,   -W = -F_SJ Finv,   U = F_SS - W F_JS   (to the parent)

which is the block LDL^T factorisation A = L D L^T with L_SJ = W, D_J = F_JJ. The top
front is the Schur complement of everything below it; it is assembled the same way, then
factored by the dense Cholesky kernels and inverted (dds_chol_factor, dds_chol_inverse).
The ordering and the tables (dds_nd_plan_build) are host code of the library, built once
per pattern.

Solve (device, every application). Forward, leaves first, one dds_nd_forward launch per
height: f = [b_J; 0] + the update vectors the children left in the front's slots (added
in child order), then g_J = Finv f_J and u = f_S - W f_J, which goes into the parent's
slots. Top: dds_nd_top_assemble, dds_dense_symv with its inverse, dds_nd_top_push.
Backward, top first, one dds_nd_backward launch per height: x_J = g_J - W^T x_S, then x
at each child's coupling set is pushed into that child's buffer. The stacked factor
[Finv; W] of a front is column-major, so the forward sweep gives each thread one row and
the backward sweep each lane one column, both streaming contiguous memory with
independent loads; on the levels near the roots, where few workgroups are in flight, the
rows' sums are split into column parts (forward) and a front's columns into narrower
items with more row parts (backward), added in part order. Every sum runs in a fixed
order.

`BatchSparseCholesky` puts the same machinery in the local-solver slot: the subdomain
matrices of a Schwarz preconditioner are factorised as one block-diagonal matrix, so one
sweep launch per height serves every subdomain.
"""

from __future__ import annotations

import ctypes
import functools
import hashlib
from typing import List

import numpy as np
import scipy.sparse as sps
import torch

from . import _lib
from .solvers import CSRBatch, SparseBatchSolver, SparseSolver

__all__ = ["SparseCholesky", "BatchSparseCholesky", "NDPlan", "factorize", "clear_plan_cache"]

# LDS capacity of the sweep kernels, in elements (f_J forward, x_S backward): 60 KiB of FP64
LDS_CAP = 7680
# largest dense top front dds_chol_inverse takes
TOP_CAP = 6144

# dds_nd_field / dds_nd_scalar codes (include/dd_solvers_hip.h)
_F = dict(SCALARS=0, PERM=1, FJ0=2, FNJ=3, FNS=4, LEVEL_PTR=5, FOFF=6, WTOFF=7, SOFF=8,
          WBASE=9, MOFF=10, MAPS=11, PUSH_DST=12, PUSH_SRC=13, PUSH_PTR=14, PARENT=15,
          CHILD_PTR=16, CHILD_IDX=17, S_IDX=18, AP_PTR=19, AP_IDX=20, AP_SRC=21, FWD_PTR=22,
          FWD_ITEMS=23, FWD_LDS=24, BWD_PTR=25, BWD_ITEMS=26, BWD_LDS=27, UOFF2=28)
_INT32_FIELDS = ("MAPS", "PUSH_SRC", "AP_IDX")
(_S_N, _S_FRONTS, _S_TOP, _S_DTOP, _S_LEVELS, _S_LOWER_LEVELS, _S_FAC_SIZE, _S_SLOT_SIZE,
 _S_XS_SIZE, _S_NNZ, _S_TREE_NODES) = range(11)


# ---------------------------------------------------------------------------------------
# Symbolic plan (native: csrc/nd_plan.cpp)
# ---------------------------------------------------------------------------------------

class NDPlan:
    """Numbering, fronts, levels, data-movement tables and work items of the solve, built
    by the library's native symbolic analysis (dds_nd_plan_build) from A's pattern.

    Fronts are numbered by height (leaves first, the top front last). Front t owns the
    unknowns [fj0[t], fj0[t] + fnj[t]) of the nested-dissection numbering (perm[i] is the
    original index of unknown i) and couples to the fns[t] unknowns S[t] of its ancestors.
    Its stacked factor [Finv (nJ x nJ); W (nS x nJ)] is stored column-major at
    fac[foff[t]:], element (r, c) at foff + c * m + r with m = nJ + nS, and W again row-major
    at fac[wtoff[t]:] for the backward sweep (the top front, nS = 0, is its symmetric dense
    inverse). Between levels data moves by pushes into fixed buffers, so no kernel follows
    an index chain: forward, a child writes its update vector into its own slot of its
    parent's front (slots[wbase[c] + maps[moff[c] + i]]); backward, a parent writes x at
    each child's coupling set into that child's buffer (xs[push_dst[k]] = value at front
    position push_src[k]). Ap is A[perm][:, perm] (stored zeros dropped, rows sorted);
    ap_src[q] is the position of its entry q in A's data after the zeros are dropped.
    """

    def __init__(self, A: sps.csr_matrix, leaf_size: int = 32, top_max: int = 4096):
        # stored zeros (the nearest-neighbour couplings of a Q1 stiffness matrix are exact
        # zeros) leave the pattern: the dissection ignores them, so the fronts must too
        A = sps.csr_matrix(A, copy=True)
        A.eliminate_zeros()
        A.sum_duplicates()  # also sorts the column indices
        n = A.shape[0]
        if A.shape != (n, n):
            raise ValueError("coarse matrix must be square")
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        h = ctypes.c_void_p()
        st = _lib._lib.dds_nd_plan_build(n, indptr.ctypes.data, indices.ctypes.data,
                                         int(leaf_size), int(top_max), ctypes.byref(h))
        if st != 0:
            msg = _lib._lib.dds_last_error().decode(errors="replace")
            raise ValueError(msg)
        try:
            f = {name: self._field(h, code, name in _INT32_FIELDS) for name, code in _F.items()}
        finally:
            _lib._lib.dds_nd_plan_free(h)
        sc = f["SCALARS"]
        self.n = n
        self.n_fronts = int(sc[_S_FRONTS])
        self.nT = int(sc[_S_TOP])
        self.d_top = int(sc[_S_DTOP])
        self.fac_size = int(sc[_S_FAC_SIZE])
        self.slot_size = int(sc[_S_SLOT_SIZE])
        self.xs_size = int(sc[_S_XS_SIZE])
        self.n_tree_nodes = int(sc[_S_TREE_NODES])
        self.perm = f["PERM"]
        self.fj0, self.fnj, self.fns = f["FJ0"], f["FNJ"], f["FNS"]
        self.m = self.fnj + self.fns
        self.level_ptr = f["LEVEL_PTR"]
        self.foff, self.wtoff, self.soff = f["FOFF"], f["WTOFF"], f["SOFF"]
        self.wbase, self.moff, self.xoff = f["WBASE"], f["MOFF"], f["MOFF"]
        self.maps = f["MAPS"]
        self.push_dst, self.push_src, self.push_ptr = f["PUSH_DST"], f["PUSH_SRC"], f["PUSH_PTR"]
        self.parent = f["PARENT"]
        self.child_ptr, self.child_idx = f["CHILD_PTR"], f["CHILD_IDX"]
        self.s_idx = f["S_IDX"]
        self.uoff2 = f["UOFF2"]
        self.ap_src = f["AP_SRC"]
        self.Ap = sps.csr_matrix((A.data[self.ap_src], f["AP_IDX"], f["AP_PTR"]), shape=(n, n))
        self.has_top = bool(self.nT)
        self.n_lower = self.n_fronts - (1 if self.has_top else 0)
        fp, fi, fl = f["FWD_PTR"], f["FWD_ITEMS"].reshape(-1, 12), f["FWD_LDS"]
        bp, bi, bl = f["BWD_PTR"], f["BWD_ITEMS"].reshape(-1, 12), f["BWD_LDS"]
        self.fwd = [(fi[fp[h]:fp[h + 1]], int(fl[h])) for h in range(fl.shape[0])]
        self.bwd = [(bi[bp[h]:bp[h + 1]], int(bl[h])) for h in range(bl.shape[0])]

    @staticmethod
    def _field(h, code: int, int32: bool) -> np.ndarray:
        n = _lib._lib.dds_nd_plan_field(h, code, None)
        if n < 0:
            raise RuntimeError(_lib._lib.dds_last_error().decode(errors="replace"))
        out = np.empty(n, dtype=np.int32 if int32 else np.int64)
        if n:
            _lib._lib.dds_nd_plan_field(h, code, out.ctypes.data)
        return out

    @property
    def n_levels(self) -> int:
        return self.level_ptr.shape[0] - 1

    # per-front lists for host-side checks (built on first use; the solve path reads the
    # flat tables)
    @functools.cached_property
    def S(self) -> List[np.ndarray]:
        """Coupling set of each front (ascending, permuted numbering)."""
        return np.split(self.s_idx, self.moff[1:-1])

    @functools.cached_property
    def children(self) -> List[List[int]]:
        """Children of each front, in slot order."""
        return [c.tolist() for c in np.split(self.child_idx, self.child_ptr[1:-1])]

    @functools.cached_property
    def cmap(self) -> List[np.ndarray]:
        """Positions of each front's coupling set in its parent's front (None for a root)."""
        return [m.astype(np.int64) if self.parent[t] >= 0 else None
                for t, m in enumerate(np.split(self.maps, self.moff[1:-1]))]

    def n_children(self, t: int) -> int:
        return int(self.child_ptr[t + 1] - self.child_ptr[t])

    def validate(self) -> None:
        """Host check of every offset the sweep and factor kernels use (before any launch)."""
        n, F = self.n, self.n_fronts
        assert np.array_equal(np.sort(self.perm), np.arange(n))
        assert self.maps.size == 0 or self.maps.min() >= 0
        assert self.push_src.size == 0 or self.push_src.min() >= 0
        assert self.push_dst.size == 0 or (self.push_dst.min() >= 0 and
                                           self.push_dst.max() < self.xs_size)
        par = self.parent
        has = par >= 0
        if np.any(has):  # every map entry lands inside the parent's front, every slot fits
            mx = np.maximum.reduceat(np.concatenate([self.maps, [-1]]).astype(np.int64),
                                     np.minimum(self.moff[:-1], self.maps.shape[0]))
            mx = np.where(self.fns > 0, mx, -1)
            assert np.all(mx[has] < self.m[par[has]])
            assert np.all(self.wbase[has] + self.m[par[has]] <= self.slot_size)
        assert np.all(self.fns[~has] == 0)
        assert np.all(self.child_idx >= 0) and np.all(self.child_idx < F)
        assert np.all(par[self.child_idx] == np.repeat(np.arange(F), np.diff(self.child_ptr)))
        assert self.s_idx.size == 0 or (self.s_idx.min() >= 0 and self.s_idx.max() < n)
        for items, lds in self.fwd:
            j0, nj, ns, foff, soff, nch, wb, mo, r0, nr, R = (items[:, k] for k in range(11))
            m = nj + ns
            assert np.all(nj >= 1) and np.all(j0 + nj <= n) and np.all(foff + m * nj <= self.fac_size)
            assert np.all(soff + nch * m <= self.slot_size)
            assert np.all(np.isin(R, (64, 128, 256)))
            assert np.all((nr >= 1) & (nr <= R)) and np.all(r0 + nr <= m)
            assert np.all((wb >= 0) | (ns == 0)) and np.all(mo + ns <= self.maps.shape[0])
            assert lds <= LDS_CAP and np.all(nj <= lds)
        for items, lds in self.bwd:
            j0, nj, ns, wt, xo, c0, nc, k0, k1, C = (items[:, k] for k in range(10))
            assert np.all(np.isin(C, (16, 32, 64)))
            assert np.all((nc >= 1) & (nc <= C)) and np.all(c0 + nc <= nj)
            assert np.all(xo + ns <= max(self.xs_size, 1)) and np.all(wt + ns * nj <= self.fac_size)
            assert np.all((0 <= k0) & (k0 <= k1) & (k1 <= self.push_src.shape[0]))
            # push sources: own columns, or the coupling set (last item of a front)
            cnt = k1 - k0
            it = np.repeat(np.arange(items.shape[0]), cnt)
            k = np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt) + np.repeat(k0, cnt)
            src = self.push_src[k].astype(np.int64)
            last = (c0 + nc == nj)[it]
            ok = ((src >= c0[it]) & (src < (c0 + nc)[it])) | \
                (last & (src >= nj[it]) & (src < (nj + ns)[it]))
            assert np.all(ok)
            assert lds <= LDS_CAP and np.all(ns <= lds)
        if self.has_top:
            t = F - 1
            k0, k1 = self.push_ptr[t], self.push_ptr[t + 1]
            assert self.push_src[k0:k1].max(initial=-1) < self.nT
            assert self.soff[t] + self.n_children(t) * self.nT <= self.slot_size
            assert self.nT <= TOP_CAP


# ---------------------------------------------------------------------------------------
# Numeric factorisation (device: dds_nd_factor per height; the top front by dds_chol_*)
# ---------------------------------------------------------------------------------------

def factorize(plan: NDPlan, values: np.ndarray, device, dtype=torch.float64) -> torch.Tensor:
    """Stacked factors [Finv; W] of every front (column-major at plan.foff) and the row-major
    copies of W (plan.wtoff), computed in FP64 on `device` and returned in `dtype`. `values`
    are the entries of plan.Ap (FP64, host or device)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("the factorisation runs on a HIP device")
    f64 = torch.float64
    I = lambda a, dt=torch.int64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    vals = torch.as_tensor(np.asarray(values, dtype=np.float64) if not torch.is_tensor(values)
                           else values, dtype=f64, device=dev).contiguous()
    fj0, fnj, fns = I(plan.fj0), I(plan.fnj), I(plan.fns)
    foff, wtoff, moff, uoff2 = I(plan.foff), I(plan.wtoff), I(plan.moff), I(plan.uoff2)
    cptr, cidx = I(plan.child_ptr), I(plan.child_idx)
    maps, s_idx = I(plan.maps, torch.int32), I(plan.s_idx)
    ap_ptr, ap_idx = I(plan.Ap.indptr), I(plan.Ap.indices, torch.int32)
    fac = torch.empty(max(plan.fac_size, 1), dtype=f64, device=dev)
    ustore = torch.empty(max(int(plan.uoff2[-1]), 1), dtype=f64, device=dev)
    info = torch.zeros(max(plan.n_fronts, 1), dtype=torch.int32, device=dev)
    s = _lib.stream()
    with torch.cuda.device(dev):
        for h in range(plan.n_levels):
            lo, hi = int(plan.level_ptr[h]), int(plan.level_ptr[h + 1])
            top = plan.has_top and h == plan.n_levels - 1
            M = plan.m[lo:hi]
            woff = np.concatenate([[0], np.cumsum(M * M)])
            work = torch.empty(max(int(woff[-1]), 1), dtype=f64, device=dev)
            _lib.call("dds_nd_factor", 1 if top else 0, lo, hi, fj0.data_ptr(), fnj.data_ptr(),
                      fns.data_ptr(), foff.data_ptr(), wtoff.data_ptr(), moff.data_ptr(),
                      uoff2.data_ptr(), cptr.data_ptr(), cidx.data_ptr(), maps.data_ptr(),
                      s_idx.data_ptr(), ap_ptr.data_ptr(), ap_idx.data_ptr(), vals.data_ptr(),
                      I(woff[:-1]).data_ptr(), work.data_ptr(), ustore.data_ptr(),
                      fac.data_ptr(), info[lo:].data_ptr(), s)
            if top:
                nT = plan.nT
                _lib.call("dds_chol_factor", _lib.F64, nT, work.data_ptr(), nT,
                          info[lo:].data_ptr(), s)
                _lib.call("dds_chol_inverse", _lib.F64, nT, work.data_ptr(), nT,
                          fac[int(plan.foff[lo]):].data_ptr(), nT, s)
            del work
        if int(info.max()) != 0:
            raise np.linalg.LinAlgError("coarse matrix is not positive definite")
    return fac if dtype == f64 else fac.to(dtype)


# ---------------------------------------------------------------------------------------
# Coarse-slot solver
# ---------------------------------------------------------------------------------------

# The last plan built, keyed by the pattern it was built for: the ordering and the tables
# depend only on the pattern of A (without stored zeros), and a run of experiments sets up
# one coarse matrix after another with the same pattern (other precisions and methods), so
# the host-side dissection of a coarse space of millions of cells is done once.
_PLAN_CACHE: dict = {}


def clear_plan_cache() -> None:
    """Drop the cached plan (its host tables outlive the solvers that used it; for a coarse
    space of millions of cells they take gigabytes of host memory)."""
    _PLAN_CACHE.clear()


def plan_for(A: sps.csr_matrix, leaf_size: int, top_max: int):
    """(plan, values): the plan of A's pattern (built, or the cached one for the same
    pattern and parameters) and A's entries in the order of the plan's permuted matrix
    plan.Ap, which `factorize` takes."""
    A = sps.csr_matrix(A, copy=True)
    A.eliminate_zeros()
    A.sum_duplicates()
    h = hashlib.blake2b(digest_size=16)
    h.update(np.ascontiguousarray(A.indptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(A.indices, dtype=np.int64).tobytes())
    key = (A.shape[0], int(leaf_size), int(top_max), h.hexdigest())
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = NDPlan(A, leaf_size, top_max)
        plan.validate()
        _PLAN_CACHE.clear()
        _PLAN_CACHE[key] = (plan, np.ascontiguousarray(A.indptr, dtype=np.int64).copy(),
                            np.ascontiguousarray(A.indices, dtype=np.int64).copy())
        return plan, plan.Ap.data
    plan, indptr, indices = plan
    if not (np.array_equal(A.indptr, indptr) and np.array_equal(A.indices, indices)):
        raise RuntimeError("cached plan does not match the pattern")  # a hash collision
    return plan, A.data[plan.ap_src]


class SparseCholesky(SparseSolver):
    """Coarse-space slot for large sparse SPD coarse matrices (any `SparseSolver` slot:
    setup(csr), solve(v) -> (x, {})); the sparse counterpart of the reference's `CUDSS()`
    (solvers.py:85-113).

    setup(A): nested-dissection ordering of A's graph and the symbolic plan on the host,
    the block LDL^T factor on the device (FP64, stored in A's dtype: float64 or float32).
    solve(b): 2 x (levels) HIP launches, every sum in a fixed order.
    """

    def __init__(self, leaf_size: int = 64, top_max: int = 4096):
        if not 0 <= top_max <= TOP_CAP:
            raise ValueError(f"top_max must be in [0, {TOP_CAP}] (got {top_max})")
        self.leaf_size = leaf_size
        self.top_max = top_max
        self.plan = None
        self.active_mode = "nested-dissection"

    def __str__(self):
        return "SparseCholesky"

    def setup(self, matrix: torch.Tensor, *args, **kwargs) -> None:
        _lib.require_device(matrix.values() if matrix.layout is torch.sparse_csr else matrix)
        if matrix.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"unsupported coarse dtype {matrix.dtype}")
        dev = matrix.device
        self.dtype = self.fac_dtype = matrix.dtype
        if matrix.layout is torch.sparse_csr:
            A = sps.csr_matrix((matrix.values().cpu().double().numpy(),
                                matrix.col_indices().cpu().numpy(),
                                matrix.crow_indices().cpu().numpy()), shape=matrix.shape)
        else:
            A = sps.csr_matrix(matrix.cpu().double().numpy())
        self._setup_host(A, dev)

    def _setup_host(self, A: sps.csr_matrix, dev) -> None:
        """Plan on the host, factor and tables on `dev` (self.dtype, the vectors' dtype,
        and self.fac_dtype, the factor's, set by the caller)."""
        self.n = A.shape[0]
        plan, values = plan_for(A, self.leaf_size, self.top_max)
        if plan.has_top and self.fac_dtype != self.dtype:
            raise ValueError("a factor stored in another dtype than the vectors needs a plan "
                             "without a top front (top_max=0)")
        self.plan = plan
        with torch.cuda.device(dev):
            self.fac = factorize(plan, values, dev, self.fac_dtype)
            I = lambda a, dt=torch.int64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt,
                                                          device=dev)
            self._maps = I(plan.maps, torch.int32)
            self._perm = I(plan.perm, torch.int32)
            # where the forward sweep reads b (perm, or a gather composed with it)
            self._perm_in = self._perm
            self._push_dst = I(plan.push_dst)
            self._push_src = I(plan.push_src, torch.int32)
            self._fwd = [(I(it.reshape(-1)), it.shape[0], lds) for it, lds in plan.fwd]
            self._bwd = [(I(it.reshape(-1)), it.shape[0], lds) for it, lds in plan.bwd]
            self._g = torch.empty(self.n, dtype=self.dtype, device=dev)
            # slot entries no child maps are read as zeros and never written
            self._slots = torch.zeros(max(plan.slot_size, 1), dtype=self.dtype, device=dev)
            self._xs = torch.zeros(max(plan.xs_size, 1), dtype=self.dtype, device=dev)
            if plan.has_top:
                self._ftop = torch.empty(plan.nT, dtype=self.dtype, device=dev)
                self._xtop = torch.empty(plan.nT, dtype=self.dtype, device=dev)

    def solve(self, rhs: torch.Tensor):
        if rhs.dtype != self.dtype:
            raise AssertionError(f"rhs dtype {rhs.dtype} != coarse dtype {self.dtype}")
        _lib.require_device(rhs)
        rhs = rhs.contiguous()
        y = torch.empty_like(rhs)
        with torch.cuda.device(rhs.device):
            self._enqueue_solve(rhs, y, None)
        return y, {}

    def _enqueue_solve(self, x, y, gate) -> None:
        dt = _lib.dtype_code(self.dtype)
        ft = _lib.dtype_code(self.fac_dtype)
        s = _lib.stream()
        g = _lib.ptr(gate)
        p = self.plan
        for items, n_items, lds in self._fwd:
            _lib.call("dds_nd_forward", dt, ft, n_items, items.data_ptr(), self._perm_in.data_ptr(),
                      self._maps.data_ptr(), self.fac.data_ptr(), x.data_ptr(),
                      self._g.data_ptr(), self._slots.data_ptr(), lds, g, s)
        if p.has_top:
            t = p.n_fronts - 1
            j0, nT = int(p.fj0[t]), p.nT
            _lib.call("dds_nd_top_assemble", dt, j0, nT, int(p.soff[t]), p.n_children(t),
                      self._perm_in.data_ptr(), x.data_ptr(), self._slots.data_ptr(),
                      self._ftop.data_ptr(), g, s)
            _lib.call("dds_dense_symv", dt, nT, self.fac[int(p.foff[t]):].data_ptr(), nT,
                      self._ftop.data_ptr(), self._xtop.data_ptr(), g, s)
            k0, k1 = int(p.push_ptr[t]), int(p.push_ptr[t + 1])
            _lib.call("dds_nd_top_push", dt, j0, nT, k1 - k0, self._perm.data_ptr(),
                      self._xtop.data_ptr(), self._push_dst[k0:].data_ptr(),
                      self._push_src[k0:].data_ptr(), self._xs.data_ptr(), y.data_ptr(), g, s)
        for items, n_items, lds in reversed(self._bwd):
            _lib.call("dds_nd_backward", dt, ft, n_items, items.data_ptr(), self._perm.data_ptr(),
                      self._push_dst.data_ptr(), self._push_src.data_ptr(), self.fac.data_ptr(),
                      self._g.data_ptr(), self._xs.data_ptr(), y.data_ptr(), lds, g, s)

    def _enqueue_solve_multi(self, x, y, st) -> None:
        raise ValueError("blocks of right-hand sides need the DenseCholesky coarse solver")

    @property
    def bytes_per_solve(self) -> int:
        e = 8 if self.dtype == torch.float64 else 4
        f = 8 if self.fac_dtype == torch.float64 else 4
        p = self.plan
        # forward: the stacked factor; backward: the row-major copy of W
        return f * int(np.sum(p.fnj * p.fnj + 2 * p.fns * p.fnj)) + 4 * e * p.n + \
            4 * e * p.xs_size

    def destroy(self):
        self.plan = None
        self.fac = None


# ---------------------------------------------------------------------------------------
# Local-solver slot: a batch of sparse SPD matrices
# ---------------------------------------------------------------------------------------

def csr_batch_to_host(batch: CSRBatch) -> sps.csr_matrix:
    """The block-diagonal matrix diag(A_0, ..., A_{n_s-1}) of a CSRBatch (FP64, host)."""
    nrows = batch.nrows.cpu().numpy().astype(np.int64)
    nnz = batch.nnz.cpu().numpy().astype(np.int64)
    ro = batch.row_offsets.cpu().numpy().astype(np.int64)
    cols = batch.col_indices.cpu().numpy().astype(np.int64)
    vals = batch.values.cpu().double().numpy()
    n_s = nrows.shape[0]
    if ro.shape[0] != int(nrows.sum()) + n_s or cols.shape[0] != int(nnz.sum()):
        raise ValueError("inconsistent CSRBatch sizes")
    rbase = np.concatenate([[0], np.cumsum(nrows)])
    ebase = np.concatenate([[0], np.cumsum(nnz)])
    sid = np.repeat(np.arange(n_s), nrows + 1)
    start = np.concatenate([[0], np.cumsum(nrows + 1)])[:-1]
    first = (np.arange(ro.shape[0]) - start[sid]) == 0
    # global row pointer: each subdomain's offsets shifted by its first entry, dropping the
    # leading zero of every subdomain but the first
    indptr = np.concatenate([[0], (ro + ebase[sid])[~first]])
    col = cols + np.repeat(rbase[:-1], nnz)
    n = int(rbase[-1])
    return sps.csr_matrix((vals, col, indptr), shape=(n, n))


class BatchSparseCholesky(SparseBatchSolver):
    """Local-solver slot for sparse subdomain matrices (`SparseBatchSolver`: setup(CSRBatch),
    solve((n_s, k)) -> flat); the counterpart of the reference's `BatchCUDSS`
    (solvers.py:390-403), which factorises every A_i with cuDSS's batch API.

    The batch is factorised as the one block-diagonal matrix diag(A_i): nested dissection
    of each subdomain's graph (its pieces are sibling subtrees, no top front), block LDL^T
    by height across all subdomains at once, and one forward and one backward sweep launch
    per height for the whole batch. The factor is computed in FP64 and stored in
    `precision` (default: the dtype of the batch values, the preconditioner precision);
    `precision=torch.float32` with FP64 values keeps FP64 vectors and sums over an FP32
    factor (half the bytes; the local solves are then exact to about 1e-7 relative).
    Inside the Schwarz preconditioners the sweeps are gated
    device launches (graph replay works), and overlapping or uneven subdomains are
    supported: the forward sweep reads the right-hand side through the subdomain index
    lists, so no gathered copy is made.
    """

    def __init__(self, leaf_size: int = 64, precision: torch.dtype | None = None):
        if precision not in (None, torch.float64, torch.float32):
            raise ValueError(f"unsupported factor precision {precision}")
        self.leaf_size = leaf_size
        self.precision = precision
        self._inner = None

    def __str__(self):
        return f"BatchSparseCholesky({self.precision or ''})"

    def setup(self, matrix_batch: CSRBatch):
        vals = matrix_batch.values
        _lib.require_device(vals)
        if vals.dtype not in (torch.float64, torch.float32):
            raise ValueError(f"unsupported local dtype {vals.dtype}")
        self.n_solvers = int(matrix_batch.nrows.shape[0])
        fac_dtype = self.precision or vals.dtype
        if vals.dtype == torch.float32 and fac_dtype == torch.float64:
            raise ValueError("an FP64 factor needs FP64 values (the preconditioner precision)")
        inner = SparseCholesky(self.leaf_size, top_max=0)
        inner.dtype, inner.fac_dtype = vals.dtype, fac_dtype
        inner._setup_host(csr_batch_to_host(matrix_batch), vals.device)
        self._inner = inner

    def _gather_from(self, sub_idx: torch.Tensor) -> None:
        """Read position p of the stacked right-hand side from x[sub_idx[p]]."""
        inner = self._inner
        inner._perm_in = sub_idx.to(torch.int64)[inner._perm.to(torch.int64)].to(torch.int32)

    def solve(self, rhs_batch: torch.Tensor) -> torch.Tensor:
        x, _ = self._inner.solve(rhs_batch.reshape(-1))
        return x

    def _enqueue_solve(self, x, y, gate) -> None:
        self._inner._enqueue_solve(x, y, gate)

    @property
    def plan(self) -> NDPlan:
        return self._inner.plan

    @property
    def bytes_per_solve(self) -> int:
        return self._inner.bytes_per_solve

    def destroy(self):
        self._inner = None
