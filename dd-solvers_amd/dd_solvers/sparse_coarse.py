"""Sparse SPD coarse solver: nested-dissection block LDL^T with level-scheduled HIP sweeps.

The reference's coarse slot is `CUDSS()` (src/dd_solvers/dd_solvers/solvers.py:85-113), a
sparse Cholesky of the coarse matrix A_C. Its experiments use coarse spaces of up to
millions of cells (one per cell of the coarse mesh), where a dense factor is out of the
question. `SparseCholesky` fills that slot without cuDSS:

Ordering (host, setup). Nested dissection of the graph of A_C by recursive bisection:
from a pseudo-peripheral node (the last node of a breadth-first search), the level set of a
second breadth-first search that splits the nodes most evenly with the fewest nodes is the
separator; the two sides are dissected in turn down to `leaf_size` nodes, and disconnected
pieces become sibling subtrees. Tree nodes at depth < d (the largest d whose separators
hold at most `top_max` nodes together) form one dense top front T, numbered last; the
subtrees below it are numbered in post-order, so every front's unknowns J are one
contiguous range and its coupling set S (the unknowns of ancestors that J touches after
elimination) lies above it.

Factorisation (device, setup). Multifrontal, one batch per height of the tree (torch FP64
dense algebra on padded fronts): the front of J is assembled from the entries of A_C in
the columns J and the children's update matrices (added child by child, so the sums are
deterministic), and

    Finv = F_JJ^-1,   W = F_SJ Finv,   U = F_SS - W F_JS   (to the parent)

which is the block LDL^T factorisation A = L D L^T with L_SJ = W, D_J = F_JJ. The top
front is the Schur complement of everything below it; its inverse is dense.

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

from dataclasses import dataclass, field
from typing import List, Tuple

import hashlib

import numpy as np
import scipy.sparse as sps
import scipy.sparse.csgraph as csgraph
import torch

from . import _lib
from .solvers import CSRBatch, SparseBatchSolver, SparseSolver

__all__ = ["SparseCholesky", "BatchSparseCholesky", "nested_dissection", "NDPlan",
           "clear_plan_cache"]

# LDS capacity of the sweep kernels, in elements (f_J forward, x_S backward): 60 KiB of FP64
LDS_CAP = 7680


# ---------------------------------------------------------------------------------------
# Ordering
# ---------------------------------------------------------------------------------------

def _bfs_levels(G: sps.csr_matrix, src: int) -> Tuple[np.ndarray, np.ndarray]:
    """Breadth-first order from src and the level of each node in that order."""
    order, pred = csgraph.breadth_first_order(G, src, directed=True, return_predecessors=True)
    nv = order.shape[0]
    pos = np.empty(G.shape[0], dtype=np.int64)
    pos[order] = np.arange(nv)
    pp = pos[pred[order[1:]]]  # parent positions, non-decreasing in breadth-first order
    bounds = [0, 1]
    while bounds[-1] < nv:
        nxt = 1 + int(np.searchsorted(pp, bounds[-1], side="left"))
        if nxt <= bounds[-1]:  # cannot happen for a breadth-first order; guard the loop
            nxt = nv
        bounds.append(nxt)
    lev = np.repeat(np.arange(len(bounds) - 1), np.diff(bounds))
    return order, lev


def _pick_level(lev: np.ndarray, balance: float = 0.47) -> int:
    """Separator level: the smallest level set leaving both sides >= `balance` of the rest,
    else the most balanced one. Near-even splits keep the tree shallow (its height is the
    number of launches of each sweep): on the 131,072-cell coarse mesh of the reference's
    S8 runs, 0.3 gives 14 heights and 0.47 gives 8 for 30% more factor bytes."""
    cnt = np.bincount(lev)
    n = lev.shape[0]
    before = np.cumsum(cnt) - cnt
    after = n - before - cnt
    cand = np.arange(1, cnt.shape[0] - 1)
    bal = np.minimum(before[cand], after[cand]) / np.maximum(n - cnt[cand], 1)
    ok = cand[bal >= balance]
    if ok.size:
        return int(ok[np.argmin(cnt[ok])])
    return int(cand[np.argmax(bal)])


@dataclass
class _Node:
    J: np.ndarray            # original indices of the unknowns eliminated here
    depth: int
    parent: int              # -1 for a root
    children: List[int] = field(default_factory=list)


def nested_dissection(G: sps.csr_matrix, leaf_size: int = 32) -> List[_Node]:
    """Dissection tree of the symmetric graph G (any values; the diagonal is ignored)."""
    G = sps.csr_matrix(G, copy=True)
    G.setdiag(0)
    G.eliminate_zeros()
    G.data[:] = 1.0
    nodes: List[_Node] = []

    def add(J, depth, parent):
        nodes.append(_Node(np.asarray(J, dtype=np.int64), depth, parent))
        if parent >= 0:
            nodes[parent].children.append(len(nodes) - 1)
        return len(nodes) - 1

    stack = [(G, np.arange(G.shape[0], dtype=np.int64), 0, -1)]
    while stack:
        H, ids, depth, parent = stack.pop()
        n = ids.shape[0]
        if n <= leaf_size:
            add(ids, depth, parent)
            continue
        order, lev = _bfs_levels(H, 0)
        if order.shape[0] < n:  # disconnected: each piece is its own subtree
            nc, lab = csgraph.connected_components(H, directed=False)
            srt = np.argsort(lab, kind="stable")
            cuts = np.concatenate([[0], np.flatnonzero(np.diff(lab[srt])) + 1, [n]])
            # one symmetric permutation, then each piece is a diagonal block (a batch of
            # thousands of subdomains costs one pass over the graph, not one per piece)
            Hs = H[srt][:, srt].tocsr()
            for a, e in reversed(list(zip(cuts[:-1], cuts[1:]))):
                stack.append((Hs[a:e, a:e], ids[srt[a:e]], depth, parent))
            continue
        order, lev = _bfs_levels(H, int(order[-1]))
        if lev[-1] < 2:  # no level splits it: keep it whole
            add(ids, depth, parent)
            continue
        L = _pick_level(lev)
        sep = order[lev == L]
        A = np.sort(order[lev < L])
        B = np.sort(order[lev > L])
        me = add(ids[np.sort(sep)], depth, parent)
        stack.append((H[B][:, B], ids[B], depth + 1, me))
        stack.append((H[A][:, A], ids[A], depth + 1, me))
    return nodes


# ---------------------------------------------------------------------------------------
# Symbolic plan
# ---------------------------------------------------------------------------------------

class NDPlan:
    """Numbering, fronts, levels, data-movement tables and work items of the solve.

    Fronts are numbered by height (leaves first, the top front last). Front t owns the
    unknowns [fj0[t], fj0[t] + fnj[t]) of the nested-dissection numbering (perm[i] is the
    original index of unknown i) and couples to the fns[t] unknowns S[t] of its ancestors.
    Its stacked factor [Finv (nJ x nJ); W (nS x nJ)] is stored column-major at
    fac[foff[t]:], element (r, c) at foff + c * m + r with m = nJ + nS, and W again row-major
    at fac[wtoff[t]:] for the backward sweep (the top front, nS = 0,
    is its symmetric dense inverse). Between levels data moves by pushes into fixed
    buffers, so no kernel follows an index chain: forward, a child writes its update
    vector into its own slot of its parent's front (slots[wbase[c] + maps[moff[c] + i]]);
    backward, a parent writes x at each child's coupling set into that child's buffer
    (xs[push_dst[k]] = value at front position push_src[k]).
    """

    def __init__(self, A: sps.csr_matrix, leaf_size: int = 32, top_max: int = 4096):
        # stored zeros (the nearest-neighbour couplings of a Q1 stiffness matrix are exact
        # zeros) leave the pattern: the dissection ignores them, so the fronts must too
        A = sps.csr_matrix(A, copy=True)
        A.eliminate_zeros()
        n = A.shape[0]
        if A.shape != (n, n):
            raise ValueError("coarse matrix must be square")
        self.n = n
        nodes = nested_dissection(A, leaf_size)
        depth = np.array([nd.depth for nd in nodes], dtype=np.int64)
        size = np.array([nd.J.shape[0] for nd in nodes], dtype=np.int64)
        dmax = int(depth.max()) + 1
        # largest d whose nodes of depth < d hold at most top_max unknowns
        d_top = 0
        while d_top < dmax and int(size[depth <= d_top].sum()) <= top_max:
            d_top += 1
        top = depth < d_top
        self.d_top = d_top
        # numbering: lower subtrees in post-order, then the top nodes
        lower_roots = [i for i, nd in enumerate(nodes)
                       if not top[i] and (nd.parent < 0 or top[nd.parent])]
        post: List[int] = []
        for r in lower_roots:
            st = [(r, False)]
            while st:
                i, done = st.pop()
                if done:
                    post.append(i)
                    continue
                st.append((i, True))
                for c in reversed(nodes[i].children):
                    st.append((c, False))
        top_nodes = [i for i in range(len(nodes)) if top[i]]
        perm = np.empty(n, dtype=np.int64)
        j0 = np.empty(len(nodes), dtype=np.int64)
        pos = 0
        for i in post + top_nodes:
            j0[i] = pos
            perm[pos:pos + size[i]] = nodes[i].J
            pos += size[i]
        assert pos == n
        self.perm = perm
        iperm = np.empty(n, dtype=np.int64)
        iperm[perm] = np.arange(n)
        Ap = A[perm][:, perm].tocsr()
        Ap.sort_indices()
        self.Ap = Ap
        nT = int(size[top].sum())
        self.nT = nT

        # fronts: lower nodes (post-order), then the top front (all top nodes merged)
        F = len(post) + (1 if nT else 0)
        fj0 = np.zeros(F, dtype=np.int64)
        fnj = np.zeros(F, dtype=np.int64)
        S: List[np.ndarray] = [None] * F
        children: List[List[int]] = [[] for _ in range(F)]
        front_of = np.full(len(nodes), -1, dtype=np.int64)
        for t, i in enumerate(post):
            front_of[i] = t
        if nT:
            for i in top_nodes:
                front_of[i] = F - 1
        for t, i in enumerate(post):
            fj0[t], fnj[t] = j0[i], size[i]
            nd = nodes[i]
            children[t] = [int(front_of[c]) for c in nd.children]
            if nd.parent >= 0 and top[nd.parent]:
                children[F - 1].append(t)
        if nT:
            fj0[F - 1], fnj[F - 1] = n - nT, nT
        indptr, indices = Ap.indptr, Ap.indices
        for t in range(len(post)):
            a, e = fj0[t], fj0[t] + fnj[t]
            adj = indices[indptr[a]:indptr[e]]
            parts = [adj[adj >= e]] + [S[c][S[c] >= e] for c in children[t]]
            S[t] = np.unique(np.concatenate(parts)).astype(np.int64)
        if nT:
            S[F - 1] = np.zeros(0, dtype=np.int64)
        fns = np.array([s.shape[0] for s in S], dtype=np.int64)
        # heights: leaves 0, parents above all their children (the top front last)
        height = np.zeros(F, dtype=np.int64)
        for t in range(F):  # children precede parents in this numbering
            if children[t]:
                height[t] = 1 + max(int(height[c]) for c in children[t])
        order = np.argsort(height, kind="stable")
        rank = np.empty(F, dtype=np.int64)
        rank[order] = np.arange(F)
        fj0, fnj, fns = fj0[order], fnj[order], fns[order]
        S = [S[i] for i in order]
        children = [[int(rank[c]) for c in children[i]] for i in order]
        height = height[order]
        self.n_fronts = F
        self.fj0, self.fnj, self.fns = fj0, fnj, fns
        self.S, self.children = S, children
        self.level_ptr = np.concatenate([[0], np.cumsum(np.bincount(height))]).astype(np.int64)
        low = slice(0, F - (1 if nT else 0))  # the top front is solved by dense kernels
        if F - (1 if nT else 0) and (int(fnj[low].max()) > LDS_CAP or int(fns[low].max()) > LDS_CAP):
            raise ValueError(f"a front exceeds {LDS_CAP} unknowns (nJ {int(fnj[low].max())}, "
                             f"nS {int(fns[low].max())}); lower leaf_size or raise top_max")

        m = fnj + fns
        self.m = m
        self.foff = np.concatenate([[0], np.cumsum(m * fnj)]).astype(np.int64)
        # W once more, row-major (element (s, c) at wtoff + s*nJ + c), for the backward sweep:
        # its lanes take consecutive columns, so this copy makes their loads coalesced
        self.wtoff = int(self.foff[-1]) + np.concatenate([[0], np.cumsum(fns * fnj)]).astype(np.int64)
        self.fac_size = int(self.wtoff[-1])
        # child maps: positions of each child's coupling set in its parent's front
        cmap: List[np.ndarray] = [None] * F
        parent = np.full(F, -1, dtype=np.int64)
        rank_in_parent = np.zeros(F, dtype=np.int64)
        for t in range(F):
            a, nj = fj0[t], fnj[t]
            for q, c in enumerate(children[t]):
                sc = S[c]
                inJ = sc < a + nj
                p = np.empty(sc.shape[0], dtype=np.int64)
                p[inJ] = sc[inJ] - a
                p[~inJ] = nj + np.searchsorted(S[t], sc[~inJ])
                assert np.all(sc[inJ] >= a)
                assert np.array_equal(S[t][p[~inJ] - nj], sc[~inJ]), "coupling set not nested"
                cmap[c] = p
                parent[c] = t
                rank_in_parent[c] = q
        self.cmap, self.parent = cmap, parent
        nch = np.array([len(c) for c in children], dtype=np.int64)
        # forward: child c adds its update vector into slot rank(c) of its parent's front,
        # slots[soff[p] + rank * m[p] + position]; positions no child maps stay zero
        self.soff = np.concatenate([[0], np.cumsum(nch * m)]).astype(np.int64)
        self.slot_size = int(self.soff[-1])
        self.wbase = np.where(parent >= 0, self.soff[:-1][np.maximum(parent, 0)] +
                              rank_in_parent * m[np.maximum(parent, 0)], -1).astype(np.int64)
        self.moff = np.concatenate([[0], np.cumsum(fns)]).astype(np.int64)
        self.maps = (np.concatenate([cmap[c] if cmap[c] is not None else np.zeros(0, np.int64)
                                     for c in range(F)]) if F else np.zeros(0)).astype(np.int32)
        # backward: x at a front's coupling set is pushed into its own buffer xs[xoff:+nS]
        # by its parent: push list of parent p = (xoff[c] + i, cmap[c][i]) sorted by source
        self.xoff = self.moff
        self.xs_size = int(self.moff[-1])
        pdst, psrc, pptr = [], [], [0]
        for t in range(F):
            d = [self.xoff[c] + np.arange(fns[c], dtype=np.int64) for c in children[t]]
            sr = [cmap[c] for c in children[t]]
            d = np.concatenate(d) if d else np.zeros(0, np.int64)
            sr = np.concatenate(sr) if sr else np.zeros(0, np.int64)
            o = np.argsort(sr, kind="stable")
            pdst.append(d[o])
            psrc.append(sr[o])
            pptr.append(pptr[-1] + d.shape[0])
        self.push_dst = np.concatenate(pdst).astype(np.int64) if F else np.zeros(0, np.int64)
        self.push_src = np.concatenate(psrc).astype(np.int32) if F else np.zeros(0, np.int32)
        self.push_ptr = np.asarray(pptr, dtype=np.int64)
        self.has_top = bool(nT)
        self.n_lower = F - (1 if nT else 0)
        self.uoff2 = np.concatenate([[0], np.cumsum(fns * fns)]).astype(np.int64)
        self._items()

    def _items(self) -> None:
        """Work items of the lower levels (the top front is solved densely).

        Forward record (12 x int64): j0, nJ, nS, foff, soff, n_children, wbase, moff, r0,
        nr, R, 0 -- rows r0 .. r0+nr-1 (nr <= R) of the front's stacked factor; the 256
        threads are R rows x 256/R column parts. Backward record: j0, nJ, nS, wtoff, xoff,
        c0, nc, k0, k1, C, 0, 0 -- columns c0 .. c0+nc-1 (nc <= C), the 256 threads being C
        columns x 256/C row parts, and the push-list entries [k0, k1) whose sources are
        those columns (the last item of a front also pushes the coupling-set values).
        Backward items take 64 columns above 32 columns, 32 above 16, else 16. On levels
        with fewer than 4096 whole-row items (too few workgroups to hide the latency of
        long per-thread sums), wide fronts split their forward sums into column parts (four
        from 32 columns on, two from 12), and long coupling sets narrow the backward items
        (at most 32 columns above 128 coupled unknowns, 16 above 256), so a thread sums at
        most about 32 rows. Measured: the leaf levels of the 64^3 Q1 boxes are fastest
        unsplit; the split levels cut the S9 coarse solve of the reference's DG runs
        (524,288 cells) by 10%."""
        self.fwd, self.bwd = [], []
        nlev = self.level_ptr.shape[0] - 1 - (1 if self.has_top else 0)
        for h in range(nlev):
            fi, bi = [], []
            lo, hi = self.level_ptr[h], self.level_ptr[h + 1]
            split = int(np.sum((self.m[lo:hi] + 255) // 256)) < 4096
            for t in range(lo, hi):
                j0, nj, ns, m = int(self.fj0[t]), int(self.fnj[t]), int(self.fns[t]), int(self.m[t])
                head = [j0, nj, ns, int(self.foff[t])]
                nch = len(self.children[t])
                R = 256 if not split else 64 if nj >= 32 else 128 if nj >= 12 else 256
                for r0 in range(0, m, R):
                    fi.append(head + [int(self.soff[t]), nch, int(self.wbase[t]),
                                      int(self.moff[t]), r0, min(R, m - r0), R, 0])
                k_lo, k_hi = int(self.push_ptr[t]), int(self.push_ptr[t + 1])
                src = self.push_src[k_lo:k_hi]
                C = 64 if nj > 32 else 32 if nj > 16 else 16
                if split:
                    C = min(C, 64 if ns <= 128 else 32 if ns <= 256 else 16)
                for c0 in range(0, nj, C):
                    nc = min(C, nj - c0)
                    k0 = k_lo + int(np.searchsorted(src, c0))
                    k1 = k_hi if c0 + nc == nj else k_lo + int(np.searchsorted(src, c0 + nc))
                    bi.append([j0, nj, ns, int(self.wtoff[t]), int(self.xoff[t]), c0, nc, k0, k1,
                               C, 0, 0])
            lf = int(self.fnj[lo:hi].max())
            lb = int(self.fns[lo:hi].max())
            self.fwd.append((np.asarray(fi, dtype=np.int64).reshape(-1, 12), lf))
            self.bwd.append((np.asarray(bi, dtype=np.int64).reshape(-1, 12), lb))

    @property
    def n_levels(self) -> int:
        return self.level_ptr.shape[0] - 1

    def validate(self) -> None:
        """Host check of every offset the sweep kernels use (before any launch)."""
        n = self.n
        assert np.array_equal(np.sort(self.perm), np.arange(n))
        assert self.maps.size == 0 or self.maps.min() >= 0
        assert self.push_src.size == 0 or self.push_src.min() >= 0
        assert self.push_dst.size == 0 or (self.push_dst.min() >= 0 and
                                           self.push_dst.max() < self.xs_size)
        for t in range(self.n_fronts):  # every map entry lands inside the parent's front
            p = self.parent[t]
            if p >= 0:
                assert self.maps[self.moff[t]:self.moff[t + 1]].max(initial=-1) < self.m[p]
                assert self.wbase[t] + self.m[p] <= self.slot_size
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
            for it in items:  # push sources: own columns, or the coupling set (last item)
                src = self.push_src[it[7]:it[8]]
                last = it[5] + it[6] == it[1]
                ok = ((src >= it[5]) & (src < it[5] + it[6])) | (last & (src >= it[1]) & (src < it[1] + it[2]))
                assert np.all(ok)
            assert lds <= LDS_CAP and np.all(ns <= lds)
        if self.has_top:
            t = self.n_fronts - 1
            k0, k1 = self.push_ptr[t], self.push_ptr[t + 1]
            assert self.push_src[k0:k1].max(initial=-1) < self.nT
            assert self.soff[t] + len(self.children[t]) * self.nT <= self.slot_size


# ---------------------------------------------------------------------------------------
# Numeric factorisation (torch, any device)
# ---------------------------------------------------------------------------------------

def _level_indices(plan: NDPlan, h: int):
    """Host index arrays of level h: padded batch shape and the scatter/gather indices."""
    lo, hi = int(plan.level_ptr[h]), int(plan.level_ptr[h + 1])
    fr = range(lo, hi)
    nj, ns = plan.fnj[lo:hi], plan.fns[lo:hi]
    Jm, Sm = int(nj.max()), int(ns.max()) if hi > lo else 0
    M = Jm + Sm
    Ap = plan.Ap
    dA, sA, dI, dU, sU = [], [], [], [], []
    pF, sF, pW, sW, pUo, sUo = [], [], [], [], [], []
    for b, t in enumerate(fr):
        a, J, s = int(plan.fj0[t]), int(plan.fnj[t]), int(plan.fns[t])
        St = plan.S[t]
        base = b * M * M
        # entries of A in the columns J (rows of the symmetric CSR), at (pos(col), pos(row))
        k0, k1 = Ap.indptr[a], Ap.indptr[a + J]
        rows = np.repeat(np.arange(a, a + J), np.diff(Ap.indptr[a:a + J + 1]))
        cols = Ap.indices[k0:k1]
        keep = cols >= a
        rows, cols, idx = rows[keep], cols[keep], np.arange(k0, k1)[keep]
        sp = np.searchsorted(St, cols)
        pc = np.where(cols < a + J, cols - a, Jm + sp)
        assert np.array_equal(St[sp[cols >= a + J]], cols[cols >= a + J])
        dA.append(base + pc * M + (rows - a))
        sA.append(idx)
        dI.append(base + np.arange(J, Jm) * (M + 1))
        # children's update matrices, child rank r in list r
        for r, c in enumerate(plan.children[t]):
            p = plan.cmap[c]
            p = np.where(p < J, p, p - J + Jm)
            ncs = p.shape[0]
            while len(dU) <= r:
                dU.append([])
                sU.append([])
            dU[r].append(base + (p[:, None] * M + p[None, :]).ravel())
            sU[r].append(plan.uoff2[c] + np.arange(ncs * ncs))
        # packing: the stacked factor [Finv (J x J); W (s x J)] column-major (element
        # (r, c) at foff + c*(J+s) + r) into fac; U (s x s) into the update store
        m = J + s
        rr, cc = np.divmod(np.arange(J * J), J)
        pF.append(plan.foff[t] + cc * m + rr)
        sF.append(b * Jm * Jm + rr * Jm + cc)
        rr, cc = np.divmod(np.arange(s * J), J)
        pW.append(np.concatenate([plan.foff[t] + cc * m + J + rr, plan.wtoff[t] + rr * J + cc]))
        sW.append(np.tile(b * Sm * Jm + rr * Jm + cc, 2))
        rr, cc = np.divmod(np.arange(s * s), max(s, 1))
        pUo.append(plan.uoff2[t] + np.arange(s * s))
        sUo.append(b * Sm * Sm + rr * Sm + cc)
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, dtype=np.int64)
    return dict(B=hi - lo, Jm=Jm, Sm=Sm, dA=cat(dA), sA=cat(sA), dI=cat(dI),
                dU=[cat(x) for x in dU], sU=[cat(x) for x in sU],
                pF=cat(pF), sF=cat(sF), pW=cat(pW), sW=cat(sW), pU=cat(pUo), sU_out=cat(sUo))


def factorize(plan: NDPlan, values: np.ndarray, device, dtype=torch.float64) -> torch.Tensor:
    """Stacked factors [Finv; W] of every front (column-major at plan.foff), in `dtype`, on
    `device`. `values` are the entries of plan.Ap (FP64)."""
    f64 = torch.float64
    vals = torch.as_tensor(np.asarray(values, dtype=np.float64), device=device)
    fac = torch.zeros(plan.fac_size, dtype=f64, device=device)
    ustore = torch.zeros(int(plan.uoff2[-1]), dtype=f64, device=device)
    T = lambda a: torch.as_tensor(a, dtype=torch.int64, device=device)
    for h in range(plan.n_levels):
        ix = _level_indices(plan, h)
        B, Jm, Sm = ix["B"], ix["Jm"], ix["Sm"]
        M = Jm + Sm
        F = torch.zeros(B * M * M, dtype=f64, device=device)
        F[T(ix["dA"])] = vals[T(ix["sA"])]
        F[T(ix["dI"])] = 1.0
        for d, s in zip(ix["dU"], ix["sU"]):  # one child rank at a time: distinct targets
            F.index_put_((T(d),), ustore[T(s)], accumulate=True)
        F = F.view(B, M, M)
        L, info = torch.linalg.cholesky_ex(F[:, :Jm, :Jm].contiguous())
        if int(info.max()) != 0:
            raise np.linalg.LinAlgError("coarse matrix is not positive definite")
        Finv = torch.cholesky_inverse(L)
        FSJ = F[:, Jm:, :Jm].contiguous()
        W = torch.matmul(FSJ, Finv)
        U = F[:, Jm:, Jm:] - torch.matmul(W, FSJ.transpose(1, 2).contiguous())
        fac[T(ix["pF"])] = Finv.reshape(-1)[T(ix["sF"])]
        fac[T(ix["pW"])] = W.reshape(-1)[T(ix["sW"])]
        ustore[T(ix["pU"])] = U.reshape(-1)[T(ix["sU_out"])]
        del F, L, Finv, FSJ, W, U
    return fac.to(dtype)


def solve_host(plan: NDPlan, fac: np.ndarray, b: np.ndarray) -> np.ndarray:
    """The device solve restated on the host over the same tables and buffers (a check of
    the plan and the factor; the device path never calls it)."""
    n = plan.n
    g = np.zeros(n)
    slots = np.zeros(max(plan.slot_size, 1))
    xs = np.zeros(max(plan.xs_size, 1))
    out = np.zeros(n)
    for items, _ in plan.fwd:                        # dds_nd_forward, level by level
        for it in items:
            j0, nj, ns, foff, soff, nch, wb, mo, r0, nr = (int(v) for v in it[:10])
            m = nj + ns
            f = b[plan.perm[j0:j0 + nj]].copy()
            for q in range(nch):
                f += slots[soff + q * m:soff + q * m + nj]
            M = fac[foff:foff + m * nj].reshape(nj, m).T   # column-major (m x nj)
            for r in range(r0, r0 + nr):
                acc = M[r] @ f
                if r < nj:
                    g[j0 + r] = acc
                else:
                    fs = sum(slots[soff + q * m + r] for q in range(nch))
                    slots[wb + plan.maps[mo + r - nj]] = fs - acc
    if plan.has_top:                                 # dds_nd_top_assemble, dense_symv, push
        t = plan.n_fronts - 1
        j0, nT = int(plan.fj0[t]), plan.nT
        ft = b[plan.perm[j0:j0 + nT]].copy()
        for q in range(len(plan.children[t])):
            ft += slots[plan.soff[t] + q * nT:plan.soff[t] + (q + 1) * nT]
        xt = fac[plan.foff[t]:plan.foff[t] + nT * nT].reshape(nT, nT) @ ft
        out[plan.perm[j0:j0 + nT]] = xt
        k0, k1 = plan.push_ptr[t], plan.push_ptr[t + 1]
        xs[plan.push_dst[k0:k1]] = xt[plan.push_src[k0:k1]]
    for items, _ in reversed(plan.bwd):              # dds_nd_backward, top level first
        for it in items:
            j0, nj, ns, wt, xo, c0, nc, k0, k1 = (int(v) for v in it[:9])
            W = fac[wt:wt + ns * nj].reshape(ns, nj)       # row-major copy of W
            xsl = xs[xo:xo + ns].copy()
            xj = g[j0 + c0:j0 + c0 + nc] - W[:, c0:c0 + nc].T @ xsl
            out[plan.perm[j0 + c0:j0 + c0 + nc]] = xj
            src = plan.push_src[k0:k1]
            val = np.where(src < nj, xj[np.clip(src - c0, 0, nc - 1)],
                           xsl[np.clip(src - nj, 0, max(ns - 1, 0))] if ns else 0.0)
            xs[plan.push_dst[k0:k1]] = val
    return out


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
    A.sort_indices()
    h = hashlib.blake2b(digest_size=16)
    h.update(np.ascontiguousarray(A.indptr, dtype=np.int64).tobytes())
    h.update(np.ascontiguousarray(A.indices, dtype=np.int64).tobytes())
    key = (A.shape[0], int(leaf_size), int(top_max), h.hexdigest())
    plan = _PLAN_CACHE.get(key)
    if plan is None:
        plan = NDPlan(A, leaf_size, top_max)
        plan.validate()
        _PLAN_CACHE.clear()
        _PLAN_CACHE[key] = plan
        return plan, plan.Ap.data
    Ap = A[plan.perm][:, plan.perm].tocsr()
    Ap.sort_indices()
    if not (np.array_equal(Ap.indptr, plan.Ap.indptr) and np.array_equal(Ap.indices, plan.Ap.indices)):
        raise RuntimeError("cached plan does not match the pattern")  # a hash collision
    return plan, Ap.data


class SparseCholesky(SparseSolver):
    """Coarse-space slot for large sparse SPD coarse matrices (any `SparseSolver` slot:
    setup(csr), solve(v) -> (x, {})); the sparse counterpart of the reference's `CUDSS()`
    (solvers.py:85-113).

    setup(A): nested-dissection ordering of A's graph and the symbolic plan on the host,
    the block LDL^T factor on the device (FP64, stored in A's dtype: float64 or float32).
    solve(b): 2 x (levels) HIP launches, every sum in a fixed order.
    """

    def __init__(self, leaf_size: int = 64, top_max: int = 4096):
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
            _lib.call("dds_nd_top_assemble", dt, j0, nT, int(p.soff[t]), len(p.children[t]),
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
