"""Host restatement of the sparse block LDL^T solver of dd_solvers.sparse_coarse (numpy,
scipy, CPU). Test infrastructure only: tests compare the library's native symbolic analysis
(csrc/nd_plan.cpp, dds_nd_plan_build) with `NDPlan` here table by table, and its device
factorisation (nd_factor_kernel, dds_nd_factor) with `factorize_host`; `solve_host` runs
the two device sweeps over the same tables. Nothing in the package imports this module.

The reference has no counterpart of its own: its coarse slot calls cuDSS 0.7.1
(src/dd_solvers/cpp/cudss.cpp:80-105, spdFactorize: cudssExecute ANALYSIS, FACTORIZATION;
:64-78 SOLVE), a sparse Cholesky behind a fill-reducing ordering. What is restated is that
published algorithm in its multifrontal block LDL^T form (the factor of a front is
[F_JJ^-1; W = F_SJ F_JJ^-1] and its Schur complement U = F_SS - W F_JS goes to the parent),
with the plan's own ordering: nested dissection by breadth-first level-set separators.

Pins (tests/test_sparse_coarse.py): factorize_host + solve_host against scipy's SuperLU
(spsolve) on 2D Laplacians, a disconnected random SPD matrix, a DG coarse matrix, a single
unknown, stored zeros and block-diagonal batches (1e-11); a factor entry planted 1e-6 off
is caught; a non-SPD matrix raises. NDPlan: perm is a permutation, every coupling set lies
above its front, an edge of A only joins fronts in an ancestor relation, children precede
parents, and validate() checks every offset the sweeps use.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np
import scipy.sparse as sps
import scipy.sparse.csgraph as csgraph

__all__ = ["nested_dissection", "NDPlan", "factorize_host", "solve_host", "LDS_CAP"]

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
        A.sum_duplicates()  # canonical rows (sorted): the searches visit neighbours in order
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


# ---------------------------------------------------------------------------------------
# Numeric factorisation
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


def factorize_host(plan, values: np.ndarray) -> np.ndarray:
    """Stacked factors [Finv; W] of every front (column-major at plan.foff) and the row-major
    copies of W (plan.wtoff), FP64, from the entries `values` of plan.Ap: per height, the
    padded fronts are factored by Cholesky (F_JJ = L L^T, Finv = L^-T L^-1), W = F_SJ Finv,
    U = F_SS - W F_SJ^T. Works with any plan object carrying the NDPlan tables."""
    vals = np.asarray(values, dtype=np.float64)
    fac = np.zeros(int(plan.fac_size))
    ustore = np.zeros(int(plan.uoff2[-1]))
    for h in range(plan.n_levels):
        ix = _level_indices(plan, h)
        B, Jm, Sm = ix["B"], ix["Jm"], ix["Sm"]
        M = Jm + Sm
        F = np.zeros(B * M * M)
        F[ix["dA"]] = vals[ix["sA"]]
        F[ix["dI"]] = 1.0
        for d, s in zip(ix["dU"], ix["sU"]):  # one child rank at a time: distinct targets
            np.add.at(F, d, ustore[s])
        F = F.reshape(B, M, M)
        FJJ = F[:, :Jm, :Jm]
        try:
            L = np.linalg.cholesky(FJJ)
        except np.linalg.LinAlgError:
            raise np.linalg.LinAlgError("coarse matrix is not positive definite") from None
        Li = np.linalg.inv(L)
        Finv = np.matmul(Li.transpose(0, 2, 1), Li)
        FSJ = F[:, Jm:, :Jm]
        W = np.matmul(FSJ, Finv)
        U = F[:, Jm:, Jm:] - np.matmul(W, FSJ.transpose(0, 2, 1))
        fac[ix["pF"]] = Finv.reshape(-1)[ix["sF"]]
        fac[ix["pW"]] = W.reshape(-1)[ix["sW"]]
        ustore[ix["pU"]] = U.reshape(-1)[ix["sU_out"]]
    return fac


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


