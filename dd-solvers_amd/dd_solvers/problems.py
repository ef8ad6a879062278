"""Synthetic model problems for the Schwarz-PCG path.

The reference assembles DG-SIPG systems with DOLFINx (src/experiments/experiments/
discretization.py:32-96). That stack is not part of this package: the workloads here are
continuous P1 triangles (2D) and Q1 hexahedra (3D) for -div(k grad u) = f on the unit
square/cube with homogeneous Dirichlet conditions. Dirichlet nodes are eliminated, so the
unknowns are the interior nodes in lexicographic order (x fastest).

The right-hand side is drawn per unknown from U(-1, 1), as the reference's `random_rhs`
does (src/experiments/experiments/experiments.py:242-246), but from a seeded generator.

Everything here is host-side numpy/scipy input generation (untimed setup). The
decomposition helpers produce the contiguous subdomain numbering that the reference's
`asm_permutation` produces (src/dd_solvers/dd_solvers/solvers.py:541-551), generalised to
uneven and overlapping subdomains.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sps

__all__ = [
    "Q1_REF_STIFFNESS_BY_DISTANCE",
    "q1_reference_stiffness",
    "assemble_q1_3d",
    "q1_assembly_operator",
    "assemble_p1_2d",
    "lognormal_coefficients",
    "uniform_rhs",
    "Decomposition",
    "box_decomposition",
    "kmeans_decomposition",
    "uniform_decomposition",
    "graph_overlap",
    "ModelProblem",
    "make_problem",
]


# ---------------------------------------------------------------------------------------
# Q1 hexahedra on a uniform grid
# ---------------------------------------------------------------------------------------

# Entry K_ab of the unit-coefficient trilinear stiffness on the unit cube depends only on
# how many coordinates of the vertices a and b differ: 1/3, 0, -1/12, -1/12.
Q1_REF_STIFFNESS_BY_DISTANCE = (1.0 / 3.0, 0.0, -1.0 / 12.0, -1.0 / 12.0)


def q1_reference_stiffness() -> np.ndarray:
    """8x8 unit-cube Q1 stiffness, vertex a = (ax, ay, az) with index ax + 2 ay + 4 az.

    Built as the tensor sum S(x)M(x)M + M(x)S(x)M + M(x)M(x)S of the 1D P1 stiffness S and
    mass M, so it is independent of the distance table above (tests compare the two).
    """
    S = np.array([[1.0, -1.0], [-1.0, 1.0]])
    M = np.array([[1.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 1.0 / 3.0]])
    # kron ordering: last factor varies fastest -> x is the last factor.
    K = np.kron(np.kron(S, M), M) + np.kron(np.kron(M, S), M) + np.kron(np.kron(M, M), S)
    return K  # index = az*4 + ay*2 + ax


def assemble_q1_3d(N: int, kappa: np.ndarray | None = None) -> sps.csr_matrix:
    """Stiffness of -div(k grad u) with Q1 elements on an N^3 grid of the unit cube.

    `kappa` has shape (N, N, N) indexed [ez, ey, ex] (one value per element); default 1.
    Unknowns are the (N-1)^3 interior nodes, x fastest. The row pattern is the full
    27-point stencil restricted to interior nodes, face couplings included (they are
    exactly zero for Q1 on a uniform grid but finite-element assembly stores them).
    """
    if kappa is None:
        kappa = np.ones((N, N, N))
    kappa = np.asarray(kappa, dtype=np.float64)
    assert kappa.shape == (N, N, N)
    m = N - 1
    h = 1.0 / N
    # Per dimension, the element slices touching node i (i = 1..N-1) and node i+d.
    sl = {0: [slice(0, m), slice(1, m + 1)], 1: [slice(1, m + 1)], -1: [slice(0, m)]}
    offsets = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    vals = np.empty((m, m, m, 27))
    valid = np.empty((m, m, m, 27), dtype=bool)
    jz, jy, jx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    for t, (dz, dy, dx) in enumerate(offsets):
        s = np.zeros((m, m, m))
        for sz in sl[dz]:
            for sy in sl[dy]:
                for sx in sl[dx]:
                    s += kappa[sz, sy, sx]
        dist = (dz != 0) + (dy != 0) + (dx != 0)
        vals[..., t] = h * Q1_REF_STIFFNESS_BY_DISTANCE[dist] * s
        valid[..., t] = (
            (jz + dz >= 0) & (jz + dz < m) & (jy + dy >= 0) & (jy + dy < m)
            & (jx + dx >= 0) & (jx + dx < m)
        )
    node = (jz * m + jy) * m + jx
    cols = np.empty((m, m, m, 27), dtype=np.int64)
    for t, (dz, dy, dx) in enumerate(offsets):
        cols[..., t] = node + (dz * m + dy) * m + dx
    valid = valid.reshape(-1, 27)
    counts = valid.sum(axis=1)
    indptr = np.zeros(m ** 3 + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    indices = cols.reshape(-1, 27)[valid].astype(np.int32)
    data = vals.reshape(-1, 27)[valid]
    A = sps.csr_matrix((data, indices, indptr), shape=(m ** 3, m ** 3))
    A.has_sorted_indices = True
    return A


def q1_assembly_operator(N: int) -> sps.csr_matrix:
    """G with A(k).data == G @ k for the CSR pattern of `assemble_q1_3d(N)`: row t of G
    holds, for stored entry t, the coefficient h * Khat(dist) of every element that
    contains both of its nodes (k per element, [ez, ey, ex] order)."""
    m = N - 1
    h = 1.0 / N
    offsets = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    jz, jy, jx = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    valid = np.empty((m, m, m, 27), dtype=bool)
    for t, (dz, dy, dx) in enumerate(offsets):
        valid[..., t] = ((jz + dz >= 0) & (jz + dz < m) & (jy + dy >= 0) & (jy + dy < m)
                         & (jx + dx >= 0) & (jx + dx < m))
    entry = np.cumsum(valid.reshape(-1)) - 1
    entry = entry.reshape(m, m, m, 27)
    # element offsets (relative to node index j, element e = j + o) touching node j and j + d
    el = {0: (0, 1), 1: (1,), -1: (0,)}
    rows, cols, vals = [], [], []
    for t, (dz, dy, dx) in enumerate(offsets):
        mask = valid[..., t]
        ent = entry[..., t][mask]
        dist = (dz != 0) + (dy != 0) + (dx != 0)
        coef = h * Q1_REF_STIFFNESS_BY_DISTANCE[dist]
        for oz in el[dz]:
            for oy in el[dy]:
                for ox in el[dx]:
                    e = ((jz + oz) * N + (jy + oy)) * N + (jx + ox)
                    rows.append(ent)
                    cols.append(e[mask])
                    vals.append(np.full(ent.shape[0], coef))
    nnz = int(valid.sum())
    G = sps.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(nnz, N ** 3))
    G.sort_indices()
    return G


def q1_element_nodes(N: int) -> np.ndarray:
    """(N^3, 8) unknown index of each element vertex (-1 on the Dirichlet boundary).

    Vertex order matches `q1_reference_stiffness` (ax + 2 ay + 4 az); elements are ordered
    [ez, ey, ex] with ex fastest.
    """
    m = N - 1
    ez, ey, ex = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    out = np.empty((N, N, N, 8), dtype=np.int64)
    for a in range(8):
        ax, ay, az = a & 1, (a >> 1) & 1, (a >> 2) & 1
        iz, iy, ix = ez + az, ey + ay, ex + ax  # node coordinates 0..N
        inside = (iz >= 1) & (iz <= m) & (iy >= 1) & (iy <= m) & (ix >= 1) & (ix <= m)
        idx = ((iz - 1) * m + (iy - 1)) * m + (ix - 1)
        out[..., a] = np.where(inside, idx, -1)
    return out.reshape(-1, 8)


# ---------------------------------------------------------------------------------------
# P1 triangles
# ---------------------------------------------------------------------------------------


def p1_mesh_2d(N: int, jitter: float = 0.0, seed: int = 0):
    """Nodes and triangles of an N x N grid on the unit square.

    Each cell (ex, ey) is split along its (0,0)-(1,1) diagonal into (n00, n10, n11) and
    (n00, n11, n01). With `jitter` > 0 interior nodes move by U(-jitter*h, jitter*h) in
    each coordinate (seeded).
    """
    h = 1.0 / N
    iy, ix = np.meshgrid(np.arange(N + 1), np.arange(N + 1), indexing="ij")
    xy = np.stack([ix * h, iy * h], axis=-1).reshape(-1, 2).astype(np.float64)
    interior = ((ix > 0) & (ix < N) & (iy > 0) & (iy < N)).reshape(-1)
    if jitter > 0:
        rng = np.random.default_rng([seed, 7])
        d = rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 2))
        xy[interior] += d
    ey, ex = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    n00 = (ey * (N + 1) + ex).reshape(-1)
    n10 = n00 + 1
    n01 = n00 + (N + 1)
    n11 = n01 + 1
    tris = np.empty((2 * N * N, 3), dtype=np.int64)
    tris[0::2] = np.stack([n00, n10, n11], axis=1)
    tris[1::2] = np.stack([n00, n11, n01], axis=1)
    return xy, tris, interior


def assemble_p1_2d(N: int, kappa: np.ndarray | None = None, jitter: float = 0.0,
                   seed: int = 0):
    """Stiffness of -div(k grad u) with P1 triangles; returns (A, xy, tris, node_to_dof).

    `kappa` has one value per triangle (default 1). Unknowns are the interior nodes in
    lexicographic order (x fastest); `node_to_dof` maps mesh nodes to unknowns (-1 on the
    boundary).
    """
    xy, tris, interior = p1_mesh_2d(N, jitter, seed)
    n_nodes = xy.shape[0]
    node_to_dof = np.full(n_nodes, -1, dtype=np.int64)
    node_to_dof[interior] = np.arange(int(interior.sum()))
    n = int(interior.sum())
    if kappa is None:
        kappa = np.ones(tris.shape[0])
    p = xy[tris]  # (T, 3, 2)
    x, y = p[..., 0], p[..., 1]
    det = (x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0])
    area = 0.5 * np.abs(det)
    # Gradients of the barycentric coordinates: grad(lambda_a) = (y_b - y_c, x_c - x_b)/det.
    bx = np.stack([y[:, 1] - y[:, 2], y[:, 2] - y[:, 0], y[:, 0] - y[:, 1]], axis=1) / det[:, None]
    by = np.stack([x[:, 2] - x[:, 1], x[:, 0] - x[:, 2], x[:, 1] - x[:, 0]], axis=1) / det[:, None]
    K = (bx[:, :, None] * bx[:, None, :] + by[:, :, None] * by[:, None, :])
    K *= (kappa * area)[:, None, None]
    dofs = node_to_dof[tris]  # (T, 3)
    rows = np.repeat(dofs, 3, axis=1).reshape(-1)
    cols = np.tile(dofs, (1, 3)).reshape(-1)
    vals = K.reshape(-1)
    keep = (rows >= 0) & (cols >= 0)
    A = sps.coo_matrix((vals[keep], (rows[keep], cols[keep])), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    A.indices = A.indices.astype(np.int32)
    return A, xy, tris, node_to_dof


# ---------------------------------------------------------------------------------------
# Coefficients and right-hand sides
# ---------------------------------------------------------------------------------------


def lognormal_coefficients(n_elements: int, seed: int, sigma: float = 1.0) -> np.ndarray:
    """k_e = exp(g_e), g_e ~ N(0, sigma^2) i.i.d. from the seeded stream [seed, 1]."""
    rng = np.random.default_rng([seed, 1])
    return np.exp(sigma * rng.standard_normal(n_elements))


def uniform_rhs(n: int, seed: int, nrhs: int | None = None) -> np.ndarray:
    """f ~ U(-1, 1) per unknown from the seeded stream [seed, 2]; shape (n,) or (n, nrhs)."""
    rng = np.random.default_rng([seed, 2])
    if nrhs is None:
        return rng.uniform(-1.0, 1.0, size=n)
    return rng.uniform(-1.0, 1.0, size=(n, nrhs))


# ---------------------------------------------------------------------------------------
# Decompositions
# ---------------------------------------------------------------------------------------


@dataclass
class Decomposition:
    """Subdomains in the contiguous numbering the Schwarz preconditioners use.

    perm            (n,)     new -> old: A_new = A[perm][:, perm], b_new = b[perm].
    owner_ptr       (n_s+1,) subdomain i owns unknowns [owner_ptr[i], owner_ptr[i+1]).
    sub_ptr         (n_s+1,) offsets into sub_idx.
    sub_idx         (sum k_i,) sorted unknowns of each subdomain (core plus overlap).
    solvers_per_coarse (n_c,) consecutive subdomains grouped into one coarse cell.
    """

    perm: np.ndarray
    owner_ptr: np.ndarray
    sub_ptr: np.ndarray
    sub_idx: np.ndarray
    solvers_per_coarse: np.ndarray = field(default=None)

    def __post_init__(self):
        self.perm = np.asarray(self.perm, dtype=np.int64)
        self.owner_ptr = np.asarray(self.owner_ptr, dtype=np.int64)
        self.sub_ptr = np.asarray(self.sub_ptr, dtype=np.int64)
        self.sub_idx = np.asarray(self.sub_idx, dtype=np.int32)
        if self.solvers_per_coarse is None:
            self.solvers_per_coarse = np.ones(self.n_subdomains, dtype=np.int64)
        self.solvers_per_coarse = np.asarray(self.solvers_per_coarse, dtype=np.int64)

    @property
    def n(self) -> int:
        return int(self.perm.shape[0])

    @property
    def n_subdomains(self) -> int:
        return int(self.owner_ptr.shape[0] - 1)

    @property
    def sizes(self) -> np.ndarray:
        return np.diff(self.sub_ptr)

    @property
    def overlapping(self) -> bool:
        return bool(self.sub_idx.shape[0] != self.n)

    def subdomain(self, i: int) -> np.ndarray:
        return self.sub_idx[self.sub_ptr[i]:self.sub_ptr[i + 1]]

    def coarse_ptr(self) -> np.ndarray:
        """(n_c+1,) unknown ranges of the coarse cells (unions of subdomain cores)."""
        scan = np.concatenate([[0], np.cumsum(self.solvers_per_coarse)])
        return self.owner_ptr[scan]


def _finish_decomposition(owner: np.ndarray, n_s: int, subsets_old, spc=None) -> Decomposition:
    """Renumber unknowns by owner (stable) and map the subdomain sets to the new numbering."""
    perm = np.argsort(owner, kind="stable")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.shape[0])
    counts = np.bincount(owner, minlength=n_s)
    owner_ptr = np.concatenate([[0], np.cumsum(counts)])
    sub_lists = [np.sort(inv[s]) for s in subsets_old]
    sub_ptr = np.concatenate([[0], np.cumsum([len(s) for s in sub_lists])])
    sub_idx = np.concatenate(sub_lists) if sub_lists else np.zeros(0, dtype=np.int64)
    return Decomposition(perm, owner_ptr, sub_ptr, sub_idx, spc)


def box_decomposition(m: int, dim: int, box: int, overlap: int) -> Decomposition:
    """Boxes of `box` interior-node layers per dimension on an m^dim lexicographic grid.

    Interior node j (0-based, per dimension) belongs to box j // box; a box extended by
    `overlap` node layers in every direction (Chebyshev distance) is the subdomain.
    Boxes are ordered lexicographically (x fastest), as are unknowns inside a box.
    """
    nb = -(-m // box)
    j = np.arange(m)
    bj = j // box
    if dim == 2:
        by, bx = np.meshgrid(bj, bj, indexing="ij")
        owner = (by * nb + bx).reshape(-1)
    else:
        bz, by, bx = np.meshgrid(bj, bj, bj, indexing="ij")
        owner = ((bz * nb + by) * nb + bx).reshape(-1)
    ranges = [(max(b * box - overlap, 0), min((b + 1) * box + overlap, m)) for b in range(nb)]
    subsets = []
    if dim == 2:
        for byi in range(nb):
            for bxi in range(nb):
                (y0, y1), (x0, x1) = ranges[byi], ranges[bxi]
                yy, xx = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
                subsets.append((yy * m + xx).reshape(-1))
    else:
        for bzi in range(nb):
            for byi in range(nb):
                for bxi in range(nb):
                    (z0, z1), (y0, y1), (x0, x1) = ranges[bzi], ranges[byi], ranges[bxi]
                    zz, yy, xx = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1),
                                             np.arange(x0, x1), indexing="ij")
                    subsets.append(((zz * m + yy) * m + xx).reshape(-1))
    return _finish_decomposition(owner, nb ** dim, subsets)


def graph_overlap(A: sps.csr_matrix, owner: np.ndarray, n_s: int, overlap: int):
    """Grow each ownership core by `overlap` layers of matrix-graph neighbours."""
    n = A.shape[0]
    P = sps.csr_matrix((np.ones(n, dtype=np.int8), (owner, np.arange(n))), shape=(n_s, n))
    G = sps.csr_matrix((np.ones(A.nnz, dtype=np.int8), A.indices, A.indptr), shape=A.shape)
    for _ in range(overlap):
        P = (P @ G).astype(bool).astype(np.int8)
    P = P.tocsr()
    P.sort_indices()
    return [P.indices[P.indptr[i]:P.indptr[i + 1]].astype(np.int64) for i in range(n_s)]


def kmeans_decomposition(A: sps.csr_matrix, xy: np.ndarray, tris: np.ndarray,
                         node_to_dof: np.ndarray, n_clusters: int, overlap: int,
                         seed: int, iters: int = 30) -> Decomposition:
    """Seeded k-means (Lloyd, Forgy start) on element centroids.

    An unknown is owned by the smallest cluster label among the triangles that contain
    it; subdomains are the cores grown by `overlap` matrix-graph layers.
    """
    from scipy.spatial import cKDTree

    cent = xy[tris].mean(axis=1)
    rng = np.random.default_rng([seed, 3])
    centers = cent[rng.choice(cent.shape[0], size=n_clusters, replace=False)].copy()
    for _ in range(iters):
        _, lab = cKDTree(centers).query(cent)
        cnt = np.bincount(lab, minlength=n_clusters)
        sx = np.bincount(lab, weights=cent[:, 0], minlength=n_clusters)
        sy = np.bincount(lab, weights=cent[:, 1], minlength=n_clusters)
        nz = cnt > 0
        centers[nz, 0] = sx[nz] / cnt[nz]
        centers[nz, 1] = sy[nz] / cnt[nz]
    _, lab = cKDTree(centers).query(cent)
    # Relabel clusters so that labels follow the centres' lexicographic (y, x) order.
    order = np.lexsort((centers[:, 0], centers[:, 1]))
    relabel = np.empty(n_clusters, dtype=np.int64)
    relabel[order] = np.arange(n_clusters)
    lab = relabel[lab]
    n = A.shape[0]
    owner = np.full(n, np.iinfo(np.int64).max, dtype=np.int64)
    d = node_to_dof[tris]
    for a in range(3):
        ok = d[:, a] >= 0
        np.minimum.at(owner, d[ok, a], lab[ok])
    present = np.unique(owner)
    remap = np.full(n_clusters, -1, dtype=np.int64)
    remap[present] = np.arange(present.shape[0])
    owner = remap[owner]
    n_s = int(present.shape[0])
    subsets = graph_overlap(A, owner, n_s, overlap)
    return _finish_decomposition(owner, n_s, subsets)


def uniform_decomposition(n: int, dofs_per_solver: int, solvers_per_coarse=None) -> Decomposition:
    """Equal-size, non-overlapping subdomains of an already-permuted system (the reference's
    drop-in case, solvers.py:554-556)."""
    assert n % dofs_per_solver == 0
    n_s = n // dofs_per_solver
    owner_ptr = np.arange(0, n + 1, dofs_per_solver)
    return Decomposition(np.arange(n), owner_ptr, owner_ptr.copy(), np.arange(n),
                         solvers_per_coarse)


# ---------------------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------------------


@dataclass
class ModelProblem:
    """A permuted system ready for the solvers: A = A0[perm][:, perm], b = b0[perm]."""

    name: str
    A: sps.csr_matrix
    b: np.ndarray
    decomposition: Decomposition
    A0: sps.csr_matrix
    b0: np.ndarray
    meta: dict


def _permute(A0: sps.csr_matrix, perm: np.ndarray) -> sps.csr_matrix:
    A = A0[perm][:, perm].tocsr()
    A.sort_indices()
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int64)
    return A


def make_problem(kind: str, N: int, box: int = 8, overlap: int = 1, seed: int = 0,
                 lognormal: bool = False, sigma: float = 1.0, nrhs: int | None = None,
                 n_clusters: int = 256, jitter: float = 0.0) -> ModelProblem:
    """Build one of the model problems.

    kind = "q1_3d": Q1 on N^3, optional lognormal k (sigma), box subdomains of `box`^3.
    kind = "p1_2d": P1 on N^2 (k = 1), box subdomains of `box`^2.
    kind = "p1_2d_kmeans": jittered P1 on N^2, `n_clusters` k-means subdomains.
    """
    meta = dict(kind=kind, N=N, box=box, overlap=overlap, seed=seed)
    if kind == "q1_3d":
        kappa = (lognormal_coefficients(N ** 3, seed, sigma).reshape(N, N, N)
                 if lognormal else None)
        A0 = assemble_q1_3d(N, kappa)
        dec = box_decomposition(N - 1, 3, box, overlap)
        meta["kappa"] = kappa
    elif kind == "p1_2d":
        A0, xy, tris, n2d = assemble_p1_2d(N)
        dec = box_decomposition(N - 1, 2, box, overlap)
    elif kind == "p1_2d_kmeans":
        A0, xy, tris, n2d = assemble_p1_2d(N, jitter=jitter, seed=seed)
        dec = kmeans_decomposition(A0, xy, tris, n2d, n_clusters, overlap, seed)
    else:
        raise ValueError(f"unknown problem kind {kind!r}")
    b0 = uniform_rhs(A0.shape[0], seed, nrhs)
    A = _permute(A0, dec.perm)
    b = b0[dec.perm]
    return ModelProblem(kind, A, np.ascontiguousarray(b), dec, A0, b0, meta)
