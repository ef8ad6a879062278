"""Discontinuous Galerkin (SIPG) systems on the reference's uniform triangle meshes.

Host-side problem generation (numpy/scipy) for the DG experiments of the reference, which
assembles them with DOLFINx (src/experiments/experiments/discretization.py:34-96) on the
mesh family `UniformMeshes(2, m)` (src/experiments/experiments/meshing.py:197-283).

Mesh S_m. The reference starts from `create_unit_square(1, 1)` (two right triangles split
by the (0,0)-(1,1) diagonal) and refines it m times with DOLFINx's Plaza refinement, every
edge marked (meshing.py:62-73, 207-213). For a right isosceles triangle with right-angle
vertex R and hypotenuse PQ that refinement bisects the hypotenuse at M = (P+Q)/2, then the
two legs at M1 = (P+R)/2 and M2 = (Q+R)/2, joining each to M: the children are
(M1; P, M), (M1; R, M), (M2; R, M), (M2; Q, M), again right isosceles with the right angle
first. So S_m is the 2^m x 2^m grid of squares with diagonals that alternate in a fixed
pattern (not the one-direction grid that midpoint subdivision would give); from level 1
on it does not depend on which diagonal S_0 used. Cells are numbered by the recursion
(the children of cell t are 4t .. 4t+3), so the S_k cell containing a cell of S_m is its
index shifted right by 2(m - k). Vertex numbering, cell numbering and the local node order
differ from DOLFINx, which changes no norm and no Schwarz iterate (the DoFs are renumbered
by `asm_permutation` anyway).

Subdomain and coarse maps. `parent_cells(k)` is the S_k cell containing each cell (the
reference composes the refinement parent maps, meshing.py:251-256); `quad_cells(k)` is the
cell of the 2^k x 2^k quadrilateral mesh C_k (meshing.py:266-283).

Space and form (discretization.py:40-96). DG_p with a nodal Lagrange basis on the points
of DOLFINx's default "DG" element (GLL-warped): the vertices, the p - 1 interior
Gauss-Lobatto-Legendre points of each edge, and interior points placed by a warp blend
(`reference_nodes`). For p <= 3 the points are fixed by symmetry (for p = 3 the interior
point is the centroid); for p >= 4 the blend places them, and the published ||b|| of the
reference's p = 5 runs pins it. rho and f are interpolated into the same space. With h_e the edge length, delta p^2 = delta * p^2 and w = rho+ rho- / (rho+ + rho-),

  a(u, v) = sum_T (rho grad u, grad v)_T
          + sum_interior (delta p^2 / h_e * w [u], [v])_e + sum_boundary (delta p^2 rho / h_e u, v)_e
          - sum_interior ([u] n+, w (grad v+ + grad v-))_e - sum_boundary (u n, rho grad v)_e
          - (the same two terms with u and v exchanged),
  b_i = (f_h, phi_i).

Every integrand is a polynomial (rho_h and f_h are piecewise polynomials; on an interior
edge the two traces of rho_h coincide, so w = rho_h / 2), and the rules used here are exact
for its degree, so the entries equal DOLFINx's up to rounding.

Pins (tests/test_dg.py): the degrees of freedom, matrix nonzeros and ||b|| that the
reference's runs record for p = 1 (S1-S11), p = 3 (S1-S10) and p = 5 (S1-S9) in
results/experiment_amgx_pcg.csv (the first residual norm of a CG run from x0 = 0 is ||b||);
the Hybrid Schwarz initial residual ||b - A T0(b)|| and the first Additive Schwarz
residuals of the reference's S10 / S8 / S8 runs (results/experiment_precisions.csv); and
second-order convergence (p = 1) to the test case's exact solution.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np
import scipy.sparse as sps
import torch

from .csr import asm_permutation

__all__ = ["UniformTriangleMesh", "UniformTetMesh", "DGTestCase", "DGProblem",
           "continuous_coefficient_2d", "constant_coefficient_2d", "continuous_coefficient_3d",
           "constant_coefficient_3d", "reference_nodes", "reference_nodes_tet", "assemble_sipg",
           "dg_experiment", "error_norm"]


# ---------------------------------------------------------------------------------------
# Mesh
# ---------------------------------------------------------------------------------------

class UniformTriangleMesh:
    """S_m: the unit square after m Plaza refinements of its two right triangles."""

    def __init__(self, m: int):
        if m < 0:
            raise ValueError("m must be >= 0")
        n = 2 ** m
        self.m, self.n = m, n
        # triangles as (right-angle vertex, hypotenuse end, hypotenuse end) on the n-grid;
        # S_0 = create_unit_square(1, 1): cells (v0, v1, v3) and (v0, v2, v3)
        T = np.array([[[n, 0], [0, 0], [n, n]], [[0, n], [0, 0], [n, n]]], dtype=np.int64)
        for _ in range(m):
            R, P, Q = T[:, 0], T[:, 1], T[:, 2]
            M, M1, M2 = (P + Q) // 2, (P + R) // 2, (Q + R) // 2
            T = np.stack([np.stack([M1, P, M], 1), np.stack([M1, R, M], 1),
                          np.stack([M2, R, M], 1), np.stack([M2, Q, M], 1)], 1).reshape(-1, 3, 2)
        j, i = np.divmod(np.arange((n + 1) ** 2), n + 1)
        self.vertices = np.stack([i, j], 1).astype(np.float64) / n
        self.cells = T[..., 1] * (n + 1) + T[..., 0]

    @property
    def num_cells(self) -> int:
        return self.cells.shape[0]

    def centroids(self) -> np.ndarray:
        return self.vertices[self.cells].mean(axis=1)

    def parent_cells(self, k: int) -> np.ndarray:
        """Index of the S_k cell that contains each cell (k <= m)."""
        if not 0 <= k <= self.m:
            raise ValueError(f"level {k} is not in [0, {self.m}]")
        return np.arange(self.num_cells, dtype=np.int64) >> (2 * (self.m - k))

    def quad_cells(self, k: int) -> np.ndarray:
        """Index of the cell of the 2^k x 2^k quadrilateral mesh C_k that contains each cell."""
        q = 2 ** k
        c = self.centroids() * q
        return np.floor(c[:, 1]).astype(np.int64) * q + np.floor(c[:, 0]).astype(np.int64)

    def level_map(self, name: str) -> np.ndarray:
        """Cell map to a mesh named as in the reference: "S<k>" or "C<k>"."""
        kind, k = name[0], int(name[1:])
        if kind == "S":
            return self.parent_cells(k)
        if kind == "C":
            return self.quad_cells(k)
        raise ValueError(f"unknown mesh id {name!r}")


# ---------------------------------------------------------------------------------------
# Test cases (src/experiments/experiments/problems.py:33-92)
# ---------------------------------------------------------------------------------------

@dataclass(frozen=True)
class DGTestCase:
    name: str
    delta: float
    rho: Callable[[np.ndarray, np.ndarray], np.ndarray]
    f: Callable[[np.ndarray, np.ndarray], np.ndarray]
    solution: Callable[[np.ndarray, np.ndarray], np.ndarray]


_pi = np.pi

constant_coefficient_2d = DGTestCase(
    "constant coefficient 2D", 7.0,
    lambda x, y: np.ones_like(x),
    lambda x, y: 2 * _pi ** 2 * np.sin(_pi * x) * np.sin(_pi * y),
    lambda x, y: np.sin(_pi * x) * np.sin(_pi * y))

continuous_coefficient_2d = DGTestCase(
    "continuous coefficient 2D", 7.0,
    lambda x, y: x ** 2 + y ** 2 + 1,
    lambda x, y: (8 * _pi) * (4 * _pi * (x ** 2 + y ** 2 + 1) * np.sin(4 * _pi * x) * np.sin(4 * _pi * y)
                              - x * np.cos(4 * _pi * x) * np.sin(4 * _pi * y)
                              - y * np.sin(4 * _pi * x) * np.cos(4 * _pi * y)),
    lambda x, y: np.sin(4 * _pi * x) * np.sin(4 * _pi * y))


# ---------------------------------------------------------------------------------------
# Reference element
# ---------------------------------------------------------------------------------------

def _gll01(n: int) -> np.ndarray:
    """The n + 1 Gauss-Lobatto-Legendre points of [0, 1] (ends and the roots of P_n')."""
    r = np.sort(np.polynomial.legendre.Legendre.basis(n).deriv().roots().real)
    return np.concatenate([[0.0], (r + 1.0) / 2.0, [1.0]])


def _gll_warp(n: int, x: np.ndarray) -> np.ndarray:
    """w(x) = sum_k (g_k - k/n) l_k(x): the displacement of the equispaced points k/n of
    [0, 1] to the GLL points g_k, interpolated by the Lagrange polynomials l_k of the
    equispaced points (zero at both ends)."""
    g = _gll01(n)
    e = np.arange(n + 1) / n
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    for k in range(1, n):
        term = np.full_like(x, g[k] - e[k])
        for m in range(n + 1):
            if m != k:
                term = term * (x - e[m]) / (e[k] - e[m])
        out = out + term
    return out


def _warped_interior(n: int) -> np.ndarray:
    """Interior points of the degree-n GLL-warped lattice of the triangle (0,0), (1,0),
    (0,1). The lattice point (x, y) = (i, j)/n, with a = 1 - x - y, moves along the two
    edge directions by the 1D warp of the barycentric differences, blended so that every
    edge gets exactly its GLL points:
        x' = x + x (a W(x - a) + y W(x - y)),   y' = y + y (a W(y - a) + x W(y - x)),
    W(t) = w(s) / (s (1 - s)) at s = (1 + t) / 2."""
    r = np.linspace(0.0, 1.0, 2 * n + 1)      # s on a half-step grid: index n + n t
    W = _gll_warp(n, r)
    W[1:-1] /= r[1:-1] * (1.0 - r[1:-1])
    pts = []
    for j in range(1, n):
        for i in range(1, n - j):
            l = n - i - j
            x, y, a = i / n, j / n, l / n
            pts.append([x + x * (a * W[n + i - l] + y * W[n + i - j]),
                        y + y * (a * W[n + j - l] + x * W[n + j - i])])
    return np.array(pts, dtype=np.float64).reshape(-1, 2)


def reference_nodes(p: int) -> np.ndarray:
    """Nodes of the DG_p basis on the triangle (0,0), (1,0), (0,1): vertices, then each
    edge's interior GLL points, then the interior points."""
    V = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    if p < 1:
        raise ValueError(f"DG degree {p} is not supported (p >= 1)")
    if p == 1:
        return V
    if p == 2:
        return np.vstack([V, [[0.5, 0.0], [0.5, 0.5], [0.0, 0.5]]])
    if p == 3:
        g = 0.5 - 0.5 / np.sqrt(5.0)  # interior GLL points of [0, 1]: g, 1 - g
        edges = []
        for a, b in ((0, 1), (1, 2), (2, 0)):
            for t in (g, 1.0 - g):
                edges.append(V[a] + t * (V[b] - V[a]))
        return np.vstack([V, np.array(edges), [[1.0 / 3.0, 1.0 / 3.0]]])
    g = _gll01(p)[1:-1]
    edges = [V[a] + t * (V[b] - V[a]) for a, b in ((0, 1), (1, 2), (2, 0)) for t in g]
    return np.vstack([V, np.array(edges), _warped_interior(p)])


def _monomials(p: int):
    return [(a, d - a) for d in range(p + 1) for a in range(d, -1, -1)]


def _mono_eval(p: int, xi: np.ndarray):
    """Monomials xi^a eta^b (a + b <= p) and their gradients at points xi (..., 2)."""
    e = _monomials(p)
    x, y = xi[..., 0], xi[..., 1]
    val = np.stack([x ** a * y ** b for a, b in e], -1)
    dx = np.stack([a * x ** max(a - 1, 0) * y ** b if a else np.zeros_like(x) for a, b in e], -1)
    dy = np.stack([b * x ** a * y ** max(b - 1, 0) if b else np.zeros_like(x) for a, b in e], -1)
    return val, np.stack([dx, dy], -1)


class _Element:
    def __init__(self, p: int):
        self.p = p
        self.nodes = reference_nodes(p)
        self.nloc = self.nodes.shape[0]
        V, _ = _mono_eval(p, self.nodes)
        self.C = np.linalg.inv(V)  # nodal basis: phi = monomials @ C

    def eval(self, xi: np.ndarray):
        """Basis values (..., nloc) and reference gradients (..., nloc, 2)."""
        m, dm = _mono_eval(self.p, xi)
        return m @ self.C, np.swapaxes(np.swapaxes(dm, -1, -2) @ self.C, -1, -2)


def _wsum(w: np.ndarray, Pi: np.ndarray, Pj: np.ndarray) -> np.ndarray:
    """(e, i, j) = sum_q w[e, q] Pi[e, q, i] Pj[e, q, j]."""
    return np.matmul(np.swapaxes(w[..., None] * Pi, 1, 2), Pj)


def _triangle_rule(degree: int):
    """Collapsed Gauss-Legendre rule on the reference triangle, exact to `degree`."""
    ng = degree // 2 + 2
    x, w = np.polynomial.legendre.leggauss(ng)
    u, wu = (x + 1) / 2, w / 2
    U, Vv = np.meshgrid(u, u, indexing="ij")
    WU, WV = np.meshgrid(wu, wu, indexing="ij")
    pts = np.stack([U * (1 - Vv), Vv], -1).reshape(-1, 2)
    wts = (WU * WV * (1 - Vv)).reshape(-1)
    return pts, wts


def _line_rule(degree: int):
    ng = degree // 2 + 1
    x, w = np.polynomial.legendre.leggauss(ng)
    return (x + 1) / 2, w / 2


# ---------------------------------------------------------------------------------------
# Assembly
# ---------------------------------------------------------------------------------------

@dataclass
class DGProblem:
    """A DG system in the mesh numbering: DoF `c * nloc + l` is node l of cell c."""

    mesh: UniformTriangleMesh
    p: int
    A: sps.csr_matrix
    b: np.ndarray
    fine_to_dofs: np.ndarray      # (num_cells, nloc)
    node_coords: np.ndarray       # (num_cells, nloc, 2)

    @property
    def block_size(self) -> int:
        return self.fine_to_dofs.shape[1]


def assemble_sipg(mesh, p: int, case: DGTestCase | None = None,
                  chunk: int = 1 << 18) -> DGProblem:
    """SIPG matrix (CSR, full block pattern) and load vector of `case` on `mesh` (a
    UniformTriangleMesh, or a UniformTetMesh for the 3D systems; the default case is the
    continuous-coefficient one of the mesh's dimension)."""
    if isinstance(mesh, UniformTetMesh):
        return _assemble_sipg_tet(mesh, p, case or continuous_coefficient_3d)
    case = case or continuous_coefficient_2d
    el = _Element(p)
    nl = el.nloc
    nc = mesh.num_cells
    X = mesh.vertices[mesh.cells]                      # (C, 3, 2)
    v0 = X[:, 0]
    Jm = np.stack([X[:, 1] - v0, X[:, 2] - v0], -1)    # columns: J[:, :, 0], J[:, :, 1]
    det = Jm[:, 0, 0] * Jm[:, 1, 1] - Jm[:, 0, 1] * Jm[:, 1, 0]
    Jinv = np.stack([np.stack([Jm[:, 1, 1], -Jm[:, 0, 1]], -1),
                     np.stack([-Jm[:, 1, 0], Jm[:, 0, 0]], -1)], -2) / det[:, None, None]
    adet = np.abs(det)
    nodes = v0[:, None, :] + np.einsum("cab,kb->cka", Jm, el.nodes)   # (C, nl, 2)
    rho_n = case.rho(nodes[..., 0], nodes[..., 1])
    f_n = case.f(nodes[..., 0], nodes[..., 1])
    dp2 = case.delta * p * p

    # Cell terms: stiffness with rho_h (degree 3p - 2) and the load (f_h, phi) (degree 2p).
    qp, qw = _triangle_rule(max(3 * p - 2, 2 * p))
    Phi, dPhi = el.eval(qp)                            # (nq, nl), (nq, nl, 2)
    nq = qw.shape[0]
    Mref = (Phi.T * qw) @ Phi
    diag = np.zeros((nc, nl, nl))
    for s in range(0, nc, chunk):
        e = slice(s, min(s + chunk, nc))
        G = np.matmul(dPhi[None], Jinv[e][:, None])    # (c, nq, nl, 2) physical gradients
        S = adet[e, None] * qw[None, :] * (rho_n[e] @ Phi.T)
        Gt = np.swapaxes(G, 1, 2).reshape(-1, nl, nq * 2)
        St = np.swapaxes(S[:, :, None, None] * G, 1, 2).reshape(-1, nl, nq * 2)
        diag[e] = np.matmul(St, np.swapaxes(Gt, 1, 2))
    b = (adet[:, None] * (f_n @ Mref)).reshape(-1)     # Mref symmetric

    # Edges: local edge l of a cell joins vertices (l, l + 1 mod 3).
    lv = np.array([[0, 1], [1, 2], [2, 0]])
    ev = mesh.cells[:, lv]                             # (C, 3, 2)
    lo, hi = ev.min(-1).reshape(-1), ev.max(-1).reshape(-1)
    key = lo * mesh.vertices.shape[0] + hi
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(ks.shape[0], bool)
    first[1:] = ks[1:] != ks[:-1]
    starts = np.flatnonzero(first)
    counts = np.diff(np.append(starts, ks.shape[0]))
    occ_cell = order // 3
    occ_edge = order % 3
    int_s = starts[counts == 2]
    bnd_s = starts[counts == 1]

    lq, lw = _line_rule(3 * p)

    def side(cells, a, b_, t):
        """Basis values, physical gradients and rho_h of `cells` at points a + t (b - a)."""
        x = a[:, None, :] + t[None, :, None] * (b_ - a)[:, None, :]
        xi = np.matmul(x - v0[cells][:, None, :], np.swapaxes(Jinv[cells], 1, 2))
        ph, dph = el.eval(xi)
        g = np.matmul(dph, Jinv[cells][:, None])
        r = np.matmul(ph, rho_n[cells][:, :, None])[..., 0]
        return ph, g, r

    def outward(cells, ledge, a, b_):
        d = b_ - a
        h = np.hypot(d[:, 0], d[:, 1])
        n = np.stack([d[:, 1], -d[:, 0]], -1) / h[:, None]
        opp = X[cells, (ledge + 2) % 3]
        flip = np.einsum("ea,ea->e", n, opp - a) > 0
        n[flip] *= -1
        return n, h

    def normal_derivative(G, n):
        return np.matmul(G, n[:, None, :, None])[..., 0]

    rows_blk, cols_blk, vals_blk = [], [], []
    diag_cells, diag_vals = [], []
    for s in range(0, int_s.shape[0], chunk):
        st = int_s[s:s + chunk]
        cp, cm = occ_cell[st], occ_cell[st + 1]
        le_p = occ_edge[st]
        a = mesh.vertices[ev[cp, le_p, 0]]
        b_ = mesh.vertices[ev[cp, le_p, 1]]
        n, h = outward(cp, le_p, a, b_)
        P1, G1, r1 = side(cp, a, b_, lq)
        P2, G2, r2 = side(cm, a, b_, lq)
        w = r1 * r2 / (r1 + r2)
        gam = dp2 / h[:, None] * w
        D1, D2 = normal_derivative(G1, n), normal_derivative(G2, n)
        W = lw[None, :] * h[:, None]
        sides = ((cp, P1, D1, 1.0), (cm, P2, D2, -1.0))
        for ci, Pi, Di, sgi in sides:
            for cj, Pj, Dj, sgj in sides:
                B = (_wsum(W * gam * (sgi * sgj), Pi, Pj) - _wsum(W * w * sgj, Di, Pj)
                     - _wsum(W * w * sgi, Pi, Dj))
                if ci is cj:
                    diag_cells.append(ci)
                    diag_vals.append(B)
                else:
                    rows_blk.append(ci)
                    cols_blk.append(cj)
                    vals_blk.append(B)
    for s in range(0, bnd_s.shape[0], chunk):
        st = bnd_s[s:s + chunk]
        c, le = occ_cell[st], occ_edge[st]
        a = mesh.vertices[ev[c, le, 0]]
        b_ = mesh.vertices[ev[c, le, 1]]
        n, h = outward(c, le, a, b_)
        P, G, r = side(c, a, b_, lq)
        D = normal_derivative(G, n)
        W = lw[None, :] * h[:, None]
        diag_cells.append(c)
        diag_vals.append(_wsum(W * dp2 * r / h[:, None], P, P) - _wsum(W * r, D, P)
                         - _wsum(W * r, P, D))
    dc = np.concatenate(diag_cells)
    diag += np.bincount((dc[:, None] * nl * nl + np.arange(nl * nl)[None, :]).reshape(-1),
                        weights=np.concatenate(diag_vals).reshape(-1),
                        minlength=nc * nl * nl).reshape(nc, nl, nl)

    # Block CSR: the diagonal block of every cell and the two couplings of every interior edge.
    br = np.concatenate([np.arange(nc)] + rows_blk)
    bc = np.concatenate([np.arange(nc)] + cols_blk)
    bv = np.concatenate([diag.reshape(nc, nl, nl)] + vals_blk)
    o = np.lexsort((bc, br))
    br, bc, bv = br[o], bc[o], bv[o]
    indptr = np.zeros(nc + 1, np.int64)
    indptr[1:] = np.cumsum(np.bincount(br, minlength=nc))
    A = sps.bsr_matrix((bv, bc, indptr), shape=(nc * nl, nc * nl)).tocsr()
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int64)
    dofs = np.arange(nc * nl, dtype=np.int64).reshape(nc, nl)
    return DGProblem(mesh, p, A, b, dofs, nodes)


def dg_experiment(problem: DGProblem, solvers_m: str, coarse_m: str):
    """The reference's experiment set-up (experiments.py:189-240) on the host: returns
    (perm, A_perm, b_perm, dofs_per_solver, solvers_per_coarse) with
    A_perm = A[perm][:, perm] (scipy CSR) in the numbering of asm_permutation, solvers on
    mesh `solvers_m` and coarse cells on mesh `coarse_m` ("S<k>" or "C<k>")."""
    mesh = problem.mesh
    f2s = mesh.level_map(solvers_m)
    if solvers_m[0] == "S":
        s2c = type(mesh)(int(solvers_m[1:])).level_map(coarse_m)
    else:
        # C_k cells map to the coarser C_j by their index on the 2^k grid
        k, j = int(solvers_m[1:]), int(coarse_m[1:])
        q = 2 ** k
        if isinstance(mesh, UniformTetMesh):
            K_, r = np.divmod(np.arange(q ** 3), q * q)
            J, I = np.divmod(r, q)
            s2c = ((K_ >> (k - j)) * (2 ** j) + (J >> (k - j))) * (2 ** j) + (I >> (k - j))
        else:
            J, I = np.divmod(np.arange(q * q), q)
            s2c = (J >> (k - j)) * (2 ** j) + (I >> (k - j))
    perm = asm_permutation(torch.as_tensor(f2s), torch.as_tensor(s2c),
                           torch.as_tensor(problem.fine_to_dofs)).numpy()
    A = problem.A[perm][:, perm].tocsr()
    A.sort_indices()
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int64)
    n_c = int(s2c.max()) + 1
    dofs_per_solver = int(np.sum(f2s == 0)) * problem.block_size
    spc = np.bincount(s2c, minlength=n_c)
    return perm, A, np.ascontiguousarray(problem.b[perm]), dofs_per_solver, spc


def error_norm(problem: DGProblem, u: np.ndarray, q: float = 2, case: DGTestCase | None = None,
               chunk: int = 1 << 18) -> float:
    """(integral |u_exact - u_h|^q dx)^(1/q) for the DG solution u (mesh numbering), by a
    collapsed Gauss rule of degree 2p + 8 per cell with the exact solution evaluated at the
    quadrature points. The reference integrates against the exact solution interpolated
    into DG_{p+2} (experiments.py:99-107, 299-305), so the two agree to that interpolation
    error."""
    if isinstance(problem.mesh, UniformTetMesh):
        return _error_norm_tet(problem, u, q, case or continuous_coefficient_3d)
    case = case or continuous_coefficient_2d
    mesh, p = problem.mesh, problem.p
    el = _Element(p)
    qp, qw = _triangle_rule(2 * p + 8)
    Phi, _ = el.eval(qp)                                   # (nq, nl)
    X = mesh.vertices[mesh.cells]
    v0 = X[:, 0]
    Jm = np.stack([X[:, 1] - v0, X[:, 2] - v0], -1)
    adet = np.abs(Jm[:, 0, 0] * Jm[:, 1, 1] - Jm[:, 0, 1] * Jm[:, 1, 0])
    U = np.asarray(u, dtype=np.float64).reshape(mesh.num_cells, el.nloc)
    total = 0.0
    for s in range(0, mesh.num_cells, chunk):
        e = slice(s, min(s + chunk, mesh.num_cells))
        pts = v0[e, None, :] + np.einsum("cab,kb->cka", Jm[e], qp)       # (c, nq, 2)
        diff = case.solution(pts[..., 0], pts[..., 1]) - U[e] @ Phi.T
        total += float(np.sum(adet[e, None] * qw[None, :] * np.abs(diff) ** q))
    return total ** (1.0 / q)


# ---------------------------------------------------------------------------------------
# 3D: the unit cube, tetrahedra (UniformMeshes(3, m), meshing.py:165-180, 197-237)
# ---------------------------------------------------------------------------------------

# The six tetrahedra of a cube (Kuhn): for an order (a, b, c) of the axes, the path
# 0 -> e_a -> e_a + e_b -> e_a + e_b + e_c through the cube's vertices. All six share the
# main diagonal of the cube.
_KUHN_ORDERS = np.array([[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]])


class UniformTetMesh:
    """S_m in 3D: the unit cube as the 2^m x 2^m x 2^m grid of cubes, each split into its
    six Kuhn tetrahedra (6 * 8^m cells, 12 * 4^m boundary faces), cell 6 c + t being order
    t of cube c = (k n + j) n + i. The reference refines create_unit_cube(1, 1, 1) (six
    tetrahedra) m times uniformly; its S_m has the same cell, face and DoF counts (the
    counts that results/benchmark_bsrmv.csv records for S5-S7). Kuhn triangulations of the
    2^k grids are nested under regular refinement, so each cell of S_m lies in one cell of
    S_k (`parent_cells`)."""

    dim = 3

    def __init__(self, m: int):
        if m < 0:
            raise ValueError("m must be >= 0")
        n = 2 ** m
        self.m, self.n = m, n
        k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        base = np.stack([i.reshape(-1), j.reshape(-1), k.reshape(-1)], 1)      # (cubes, 3)
        eye = np.eye(3, dtype=np.int64)
        path = np.zeros((6, 4, 3), dtype=np.int64)
        for t, (a, b_, c) in enumerate(_KUHN_ORDERS):
            path[t, 1] = eye[a]
            path[t, 2] = eye[a] + eye[b_]
            path[t, 3] = 1
        P = base[:, None, None, :] + path[None]                                # (cubes, 6, 4, 3)
        self.cells = ((P[..., 2] * (n + 1) + P[..., 1]) * (n + 1) + P[..., 0]).reshape(-1, 4)
        kk, jj, ii = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
        self.vertices = np.stack([ii.reshape(-1), jj.reshape(-1), kk.reshape(-1)], 1).astype(np.float64) / n

    @property
    def num_cells(self) -> int:
        return self.cells.shape[0]

    def centroids(self) -> np.ndarray:
        return self.vertices[self.cells].mean(axis=1)

    def parent_cells(self, k: int) -> np.ndarray:
        """Index of the S_k cell that contains each cell (k <= m): its cube on the 2^k grid,
        and the Kuhn tetrahedron of that cube holding the centroid (the axis order sorting
        the centroid's local coordinates in decreasing order)."""
        if not 0 <= k <= self.m:
            raise ValueError(f"level {k} is not in [0, {self.m}]")
        q = 2 ** k
        c = self.centroids() * q
        cube = np.minimum(np.floor(c).astype(np.int64), q - 1)
        u = c - cube
        order = np.argsort(-u, axis=1, kind="stable")
        code = order[:, 0] * 9 + order[:, 1] * 3 + order[:, 2]
        lut = np.full(27, -1, dtype=np.int64)
        lut[_KUHN_ORDERS[:, 0] * 9 + _KUHN_ORDERS[:, 1] * 3 + _KUHN_ORDERS[:, 2]] = np.arange(6)
        cid = (cube[:, 2] * q + cube[:, 1]) * q + cube[:, 0]
        return cid * 6 + lut[code]

    def quad_cells(self, k: int) -> np.ndarray:
        """Index of the cell of the 2^k x 2^k x 2^k hexahedral mesh C_k holding each cell."""
        q = 2 ** k
        c = np.minimum(np.floor(self.centroids() * q).astype(np.int64), q - 1)
        return (c[:, 2] * q + c[:, 1]) * q + c[:, 0]

    def level_map(self, name: str) -> np.ndarray:
        kind, k = name[0], int(name[1:])
        if kind == "S":
            return self.parent_cells(k)
        if kind == "C":
            return self.quad_cells(k)
        raise ValueError(f"unknown mesh id {name!r}")

    def faces(self):
        """Face incidences: local face l of a cell is the one opposite its vertex l. Returns
        (cell, local face) of the interior faces' two sides (arrays of the same length,
        first side the lower index in sorted order) and of the boundary faces."""
        lf = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
        # a face is a chain w, w + a, w + a + b of grid vertices (a, b disjoint 0/1 steps), so
        # the sum of its vertices' grid coordinates, 3 w + 2 a + b, identifies it
        n = self.n
        g = np.rint(self.vertices * n).astype(np.int64)
        S = g[self.cells[:, lf]].sum(axis=2).reshape(-1, 3)          # (4 C, 3) in [0, 3n]
        key = (S[:, 2] * (3 * n + 1) + S[:, 1]) * (3 * n + 1) + S[:, 0]
        order = np.argsort(key, kind="stable")
        ks = key[order]
        first = np.ones(ks.shape[0], bool)
        first[1:] = ks[1:] != ks[:-1]
        starts = np.flatnonzero(first)
        counts = np.diff(np.append(starts, ks.shape[0]))
        it, bd = starts[counts == 2], starts[counts == 1]
        return ((order[it] // 4, order[it] % 4), (order[it + 1] // 4, order[it + 1] % 4),
                (order[bd] // 4, order[bd] % 4))


constant_coefficient_3d = DGTestCase(
    "constant coefficient 3D", 17.0,
    lambda x, y, z: np.ones_like(x),
    lambda x, y, z: 3 * _pi ** 2 * np.sin(_pi * x) * np.sin(_pi * y) * np.sin(_pi * z),
    lambda x, y, z: np.sin(_pi * x) * np.sin(_pi * y) * np.sin(_pi * z))

continuous_coefficient_3d = DGTestCase(
    "continuous coefficient 3D", 17.0,
    lambda x, y, z: x ** 2 + y ** 2 + z ** 2 + 1,
    lambda x, y, z: (8 * _pi) * (
        6 * _pi * (x ** 2 + y ** 2 + z ** 2 + 1)
        * np.sin(4 * _pi * x) * np.sin(4 * _pi * y) * np.sin(4 * _pi * z)
        - x * np.cos(4 * _pi * x) * np.sin(4 * _pi * y) * np.sin(4 * _pi * z)
        - y * np.sin(4 * _pi * x) * np.cos(4 * _pi * y) * np.sin(4 * _pi * z)
        - z * np.sin(4 * _pi * x) * np.sin(4 * _pi * y) * np.cos(4 * _pi * z)),
    lambda x, y, z: np.sin(4 * _pi * x) * np.sin(4 * _pi * y) * np.sin(4 * _pi * z))


def reference_nodes_tet(p: int) -> np.ndarray:
    """Nodes of the DG_p basis on the tetrahedron (0,0,0), (1,0,0), (0,1,0), (0,0,1), p <= 3:
    vertices, each edge's interior GLL points, and for p = 3 each face's centroid (the
    points of DOLFINx's default DG element, fixed by symmetry up to p = 3)."""
    V = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    if not 1 <= p <= 3:
        raise ValueError(f"3D DG degree {p} is not supported (1 <= p <= 3)")
    if p == 1:
        return V
    edges = ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3))
    g = _gll01(p)[1:-1]
    pts = [V] + [np.array([V[a] + t * (V[b] - V[a]) for t in g]) for a, b in edges]
    if p == 3:
        pts.append(np.array([V[list(f)].mean(axis=0) for f in ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))]))
    return np.vstack(pts)


def _mono3(p: int):
    return [(a, b, d - a - b) for d in range(p + 1) for a in range(d, -1, -1) for b in range(d - a, -1, -1)]


def _mono3_eval(p: int, xi: np.ndarray):
    """Monomials x^a y^b z^c (a + b + c <= p) and their gradients at points xi (..., 3)."""
    e = _mono3(p)
    x, y, z = xi[..., 0], xi[..., 1], xi[..., 2]

    def pw(t, k):
        return t ** k if k > 0 else np.ones_like(t)

    val = np.stack([pw(x, a) * pw(y, b) * pw(z, c) for a, b, c in e], -1)
    dx = np.stack([a * pw(x, a - 1) * pw(y, b) * pw(z, c) if a else np.zeros_like(x) for a, b, c in e], -1)
    dy = np.stack([b * pw(x, a) * pw(y, b - 1) * pw(z, c) if b else np.zeros_like(x) for a, b, c in e], -1)
    dz = np.stack([c * pw(x, a) * pw(y, b) * pw(z, c - 1) if c else np.zeros_like(x) for a, b, c in e], -1)
    return val, np.stack([dx, dy, dz], -1)


class _TetElement:
    def __init__(self, p: int):
        self.p = p
        self.nodes = reference_nodes_tet(p)
        self.nloc = self.nodes.shape[0]
        V, _ = _mono3_eval(p, self.nodes)
        self.C = np.linalg.inv(V)

    def eval(self, xi: np.ndarray):
        m, dm = _mono3_eval(self.p, xi)
        return m @ self.C, np.swapaxes(np.swapaxes(dm, -1, -2) @ self.C, -1, -2)


def _tet_rule(degree: int):
    """Collapsed (Duffy) Gauss-Legendre rule on the reference tetrahedron, exact to `degree`:
    x = u (1 - v)(1 - w), y = v (1 - w), z = w, Jacobian (1 - v)(1 - w)^2."""
    ng = degree // 2 + 2
    t, wt = np.polynomial.legendre.leggauss(ng)
    t, wt = (t + 1) / 2, wt / 2
    U, V, W = np.meshgrid(t, t, t, indexing="ij")
    WU, WV, WW = np.meshgrid(wt, wt, wt, indexing="ij")
    pts = np.stack([U * (1 - V) * (1 - W), V * (1 - W), W], -1).reshape(-1, 3)
    wts = (WU * WV * WW * (1 - V) * (1 - W) ** 2).reshape(-1)
    return pts, wts


def _assemble_sipg_tet(mesh: UniformTetMesh, p: int, case: DGTestCase, chunk: int = 1 << 16) -> DGProblem:
    """The SIPG form of assemble_sipg on tetrahedra: faces in place of edges, the penalty
    with h_e the face's longest edge (ufl.MaxFacetEdgeLength, discretization.py:49-50)."""
    el = _TetElement(p)
    nl = el.nloc
    nc = mesh.num_cells
    X = mesh.vertices[mesh.cells]                          # (C, 4, 3)
    v0 = X[:, 0]
    Jm = np.stack([X[:, 1] - v0, X[:, 2] - v0, X[:, 3] - v0], -1)   # columns are the edges
    det = np.linalg.det(Jm)
    Jinv = np.linalg.inv(Jm)
    adet = np.abs(det)
    nodes = v0[:, None, :] + np.einsum("cab,kb->cka", Jm, el.nodes)  # (C, nl, 3)
    rho_n = case.rho(nodes[..., 0], nodes[..., 1], nodes[..., 2])
    f_n = case.f(nodes[..., 0], nodes[..., 1], nodes[..., 2])
    dp2 = case.delta * p * p

    qp, qw = _tet_rule(max(3 * p - 2, 2 * p))
    Phi, dPhi = el.eval(qp)
    nq = qw.shape[0]
    Mref = (Phi.T * qw) @ Phi
    diag = np.zeros((nc, nl, nl))
    for s in range(0, nc, chunk):
        e = slice(s, min(s + chunk, nc))
        G = np.matmul(dPhi[None], Jinv[e][:, None])        # (c, nq, nl, 3)
        S = adet[e, None] * qw[None, :] * (rho_n[e] @ Phi.T)
        Gt = np.swapaxes(G, 1, 2).reshape(-1, nl, nq * 3)
        St = np.swapaxes(S[:, :, None, None] * G, 1, 2).reshape(-1, nl, nq * 3)
        diag[e] = np.matmul(St, np.swapaxes(Gt, 1, 2))
    b = (adet[:, None] * (f_n @ Mref)).reshape(-1)

    lf = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    (cp_all, lp_all), (cm_all, _), (cb_all, lb_all) = mesh.faces()
    fq, fw = _triangle_rule(3 * p)                         # on the face, weights sum to 1/2

    def face_geometry(cells, lface):
        F = X[cells[:, None], lf[lface]]                    # (e, 3, 3) the face's vertices
        e1, e2 = F[:, 1] - F[:, 0], F[:, 2] - F[:, 0]
        nrm = np.cross(e1, e2)
        area2 = np.linalg.norm(nrm, axis=1)
        n = nrm / area2[:, None]
        opp = X[cells, lface]
        flip = np.einsum("ea,ea->e", n, opp - F[:, 0]) > 0
        n[flip] *= -1
        he = np.max(np.stack([np.linalg.norm(e1, axis=1), np.linalg.norm(e2, axis=1),
                              np.linalg.norm(F[:, 2] - F[:, 1], axis=1)], 1), axis=1)
        x = F[:, None, 0] + fq[None, :, 0, None] * e1[:, None] + fq[None, :, 1, None] * e2[:, None]
        W = fw[None, :] * area2[:, None]                   # 2 x area x (weights of area 1/2)
        return n, he, x, W

    def side(cells, x):
        xi = np.matmul(x - v0[cells][:, None, :], np.swapaxes(Jinv[cells], 1, 2))
        ph, dph = el.eval(xi)
        g = np.matmul(dph, Jinv[cells][:, None])
        r = np.matmul(ph, rho_n[cells][:, :, None])[..., 0]
        return ph, g, r

    def normal_derivative(G, n):
        return np.matmul(G, n[:, None, :, None])[..., 0]

    rows_blk, cols_blk, vals_blk = [], [], []
    diag_cells, diag_vals = [], []
    for s in range(0, cp_all.shape[0], chunk):
        cp, lp, cm = cp_all[s:s + chunk], lp_all[s:s + chunk], cm_all[s:s + chunk]
        n, he, x, W = face_geometry(cp, lp)
        P1, G1, r1 = side(cp, x)
        P2, G2, r2 = side(cm, x)
        w = r1 * r2 / (r1 + r2)
        gam = dp2 / he[:, None] * w
        D1, D2 = normal_derivative(G1, n), normal_derivative(G2, n)
        sides = ((cp, P1, D1, 1.0), (cm, P2, D2, -1.0))
        for ci, Pi, Di, sgi in sides:
            for cj, Pj, Dj, sgj in sides:
                B = (_wsum(W * gam * (sgi * sgj), Pi, Pj) - _wsum(W * w * sgj, Di, Pj)
                     - _wsum(W * w * sgi, Pi, Dj))
                if ci is cj:
                    diag_cells.append(ci)
                    diag_vals.append(B)
                else:
                    rows_blk.append(ci)
                    cols_blk.append(cj)
                    vals_blk.append(B)
    for s in range(0, cb_all.shape[0], chunk):
        c, lfc = cb_all[s:s + chunk], lb_all[s:s + chunk]
        n, he, x, W = face_geometry(c, lfc)
        P, G, r = side(c, x)
        D = normal_derivative(G, n)
        diag_cells.append(c)
        diag_vals.append(_wsum(W * dp2 * r / he[:, None], P, P) - _wsum(W * r, D, P)
                         - _wsum(W * r, P, D))
    dc = np.concatenate(diag_cells)
    diag += np.bincount((dc[:, None] * nl * nl + np.arange(nl * nl)[None, :]).reshape(-1),
                        weights=np.concatenate(diag_vals).reshape(-1),
                        minlength=nc * nl * nl).reshape(nc, nl, nl)

    br = np.concatenate([np.arange(nc)] + rows_blk)
    bc = np.concatenate([np.arange(nc)] + cols_blk)
    bv = np.concatenate([diag] + vals_blk)
    o = np.lexsort((bc, br))
    br, bc, bv = br[o], bc[o], bv[o]
    indptr = np.zeros(nc + 1, np.int64)
    indptr[1:] = np.cumsum(np.bincount(br, minlength=nc))
    A = sps.bsr_matrix((bv, bc, indptr), shape=(nc * nl, nc * nl)).tocsr()
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int64)
    dofs = np.arange(nc * nl, dtype=np.int64).reshape(nc, nl)
    return DGProblem(mesh, p, A, b, dofs, nodes)


def _error_norm_tet(problem: DGProblem, u: np.ndarray, q: float, case: DGTestCase,
                    chunk: int = 1 << 16) -> float:
    mesh, p = problem.mesh, problem.p
    el = _TetElement(p)
    qp, qw = _tet_rule(2 * p + 8)
    Phi, _ = el.eval(qp)
    X = mesh.vertices[mesh.cells]
    v0 = X[:, 0]
    Jm = np.stack([X[:, 1] - v0, X[:, 2] - v0, X[:, 3] - v0], -1)
    adet = np.abs(np.linalg.det(Jm))
    U = np.asarray(u, dtype=np.float64).reshape(mesh.num_cells, el.nloc)
    total = 0.0
    for s in range(0, mesh.num_cells, chunk):
        e = slice(s, min(s + chunk, mesh.num_cells))
        pts = v0[e, None, :] + np.einsum("cab,kb->cka", Jm[e], qp)
        diff = case.solution(pts[..., 0], pts[..., 1], pts[..., 2]) - U[e] @ Phi.T
        total += float(np.sum(adet[e, None] * qw[None, :] * np.abs(diff) ** q))
    return total ** (1.0 / q)
