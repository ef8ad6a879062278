"""DG-SIPG systems of the reference's experiments (dd_solvers.dg) against the reference's
published results (tests/golden/dg_published.json, extracted from its results/*.csv by
tests/golden/make_dg_published.py), the SIPG convergence order, and the oracle."""

import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.pcg import cg
from oracle.schwarz import BlockSchwarzOracle, SchwarzOracle

from dd_solvers import dg

PUB = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dg_published.json")))


def _nloc(p):
    return (p + 1) * (p + 2) // 2


@pytest.mark.parametrize("p,level", [(1, m) for m in range(1, 9)] + [(3, m) for m in range(1, 7)]
                         + [(5, m) for m in range(1, 6)])
def test_published_dofs_nnz_and_load_norm(p, level):
    """DoFs, nonzeros and ||b|| of the reference's DOLFINx assembly, level by level. For
    p = 5, ||b|| depends on where the six interior nodes of the element sit, so it pins the
    warp blend of reference_nodes."""
    ref = PUB["levels"][f"p{p}_S{level}"]
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(level), p)
    assert pr.A.shape[0] == ref["dofs"]
    assert pr.A.nnz == ref["nnz"]
    nb = float(np.linalg.norm(pr.b))
    if ref["b_norm"] < 1e-12:   # S1, S2 at p = 1: f vanishes at every vertex
        assert nb < 1e-12
    else:
        assert abs(nb - ref["b_norm"]) <= 1e-13 * ref["b_norm"], (nb, ref["b_norm"])
    assert abs(pr.A - pr.A.T).max() <= 1e-14 * abs(pr.A).max()


def test_published_sizes_of_the_large_levels():
    """Cells and interior edges of our S_m give the published DoFs and nonzeros of every
    level the reference ran (nnz = nloc^2 (cells + 2 interior edges))."""
    for key, ref in PUB["levels"].items():
        if ref["level"] > 10:
            continue
        mesh = dg.UniformTriangleMesh(ref["level"])
        e = np.sort(mesh.cells[:, [[0, 1], [1, 2], [2, 0]]], axis=-1).reshape(-1, 2)
        key_ = e[:, 0] * (mesh.n + 1) ** 2 + e[:, 1]
        _, counts = np.unique(key_, return_counts=True)
        nl = _nloc(ref["p"])
        assert nl * mesh.num_cells == ref["dofs"], key
        assert nl * nl * (mesh.num_cells + 2 * int((counts == 2).sum())) == ref["nnz"], key


def test_mesh_hierarchy():
    """Children lie inside their parents; S_k and C_k maps agree with point location."""
    fine = dg.UniformTriangleMesh(5)
    c = fine.centroids()
    for k in range(0, 5):
        par = fine.parent_cells(k)
        coarse = dg.UniformTriangleMesh(k)
        V = coarse.vertices[coarse.cells[par]]          # the parent triangle of each cell
        d = lambda a, b, p_: (b[:, 0] - a[:, 0]) * (p_[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (p_[:, 0] - a[:, 0])
        s = np.stack([d(V[:, 0], V[:, 1], c), d(V[:, 1], V[:, 2], c), d(V[:, 2], V[:, 0], c)], 1)
        assert np.all((s > 0).all(1) | (s < 0).all(1)), k
        assert np.array_equal(np.bincount(par), np.full(coarse.num_cells, 4 ** (5 - k)))
    q = fine.quad_cells(3)
    assert np.array_equal(q, np.floor(c[:, 1] * 8).astype(int) * 8 + np.floor(c[:, 0] * 8).astype(int))


@pytest.mark.parametrize("p,levels,order", [(1, (5, 6, 7), 2), (2, (3, 4, 5), 3), (3, (3, 4, 5), 4)])
def test_sipg_converges_at_the_expected_order(p, levels, order):
    """Nodal RMS error against sin(4 pi x) sin(4 pi y) falls as h^(p+1)."""
    errs = []
    for m in levels:
        pr = dg.assemble_sipg(dg.UniformTriangleMesh(m), p)
        u = spla.spsolve(pr.A.tocsc(), pr.b)
        xy = pr.node_coords.reshape(-1, 2)
        ex = dg.continuous_coefficient_2d.solution(xy[:, 0], xy[:, 1])
        errs.append(float(np.sqrt(np.mean((u - ex) ** 2))))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > order - 0.3, (errs, rates)


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
@pytest.mark.parametrize("solvers,coarse", [("S2", "S1"), ("S3", "S3"), ("S3", "C1")])
def test_block_oracle_matches_general_oracle(kind, solvers, coarse):
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(4), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, solvers, coarse)
    n_s = A.shape[0] // k
    ptr = np.arange(n_s + 1) * k
    gen = SchwarzOracle(A, ptr, ptr, np.arange(A.shape[0]), spc, kind=kind)
    blk = BlockSchwarzOracle(A, k, spc, kind=kind)
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    ref = gen.apply(x)
    assert np.abs(blk.apply(x) - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(blk.T0(b), gen.T0(b), rtol=1e-12, atol=1e-14 * np.abs(b).max())


@pytest.fixture(scope="module")
def s10_s8_s8():
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(10), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S8", "S8")
    return A, b, BlockSchwarzOracle(A, k, spc)


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_reproduces_published_fp64_runs_s10_s8_s8(s10_s8_s8, kind):
    """The reference's CG(AdditiveSchwarz / HybridSchwarz(float64, Inv(), CUDSS)) runs on
    the continuous-coefficient problem, p = 1, S10 fine mesh (6,291,456 DoFs), S8
    subdomains and coarse cells (131,072 each), rtol 1e-9: iteration count and residual
    history (results/experiment_precisions.csv)."""
    A, b, M = s10_s8_s8
    M.kind = kind
    ref = PUB["runs"][f"p1_S10_S8_S8_{kind}"]
    x, md = cg(lambda v: A @ v, b, precond=M.apply, T0=M.T0 if kind == "hybrid" else None,
               rtol=ref["rtol"], maxiter=2000)
    assert md["iterations"] == ref["iterations"]
    got, pub = np.array(md["residual norms"]), np.array(ref["residual_norms"])
    rel = np.abs(got - pub) / pub
    assert rel[:10].max() < 1e-11, rel[:10]
    assert rel.max() < 1e-3, rel.max()


def test_reference_nodes_gll_on_edges_and_planted_shift():
    """Every degree's nodes: GLL points on each edge, interior points strictly inside and
    symmetric under the triangle's reflection x <-> y; p = 3 gives the centroid. A planted
    shift of one interior p = 5 node moves ||b|| of S2 far off the published value."""
    from dd_solvers.dg import _gll01
    for p in range(1, 7):
        X = dg.reference_nodes(p)
        assert X.shape == (_nloc(p), 2)
        g = _gll01(p)[1:-1]
        e0 = X[3:3 + (p - 1)]
        assert np.allclose(e0[:, 0], g) and np.allclose(e0[:, 1], 0.0)
        inner = X[3 + 3 * (p - 1):]
        assert np.all(inner > 1e-3) and np.all(inner.sum(1) < 1 - 1e-3)
        mirrored = np.sort(inner[:, ::-1].round(12), axis=0)
        assert np.allclose(np.sort(inner.round(12), axis=0), mirrored)
    assert np.allclose(dg.reference_nodes(3)[-1], [1 / 3, 1 / 3])
    ref = PUB["levels"]["p5_S2"]["b_norm"]
    orig = dg.reference_nodes
    try:
        def shifted(p):
            X = orig(p).copy()
            if p == 5:
                X[-1] += [1e-3, -1e-3]
            return X
        dg.reference_nodes = shifted
        nb = float(np.linalg.norm(dg.assemble_sipg(dg.UniformTriangleMesh(2), 5).b))
    finally:
        dg.reference_nodes = orig
    assert abs(nb - ref) > 1e-8 * ref


# ---------------------------------------------------------------------------------------
# 3D: tetrahedral S_m (the reference's UniformMeshes(3, m)), DG_p for p <= 3
# ---------------------------------------------------------------------------------------

def _nloc3(p):
    return (p + 1) * (p + 2) * (p + 3) // 6


def test_tet_mesh_counts_match_the_published_bsr_shapes():
    """6 * 8^m cells and 12 * 4^m boundary faces on the explicit meshes S0-S5, hence a block
    per cell and two per interior face: nnzb = T + (4 T - 12 * 4^m). The reference's three
    3D rows of its BSR benchmark (S7 p = 1, S6 p = 2, S5 p = 3) give exactly these DoFs,
    block counts and nonzeros (S5 from the explicit mesh, S6 and S7 from the counts)."""
    for m in range(6):
        M = dg.UniformTetMesh(m)
        (cp, _), (cm, _), (cb, _) = M.faces()
        assert M.num_cells == 6 * 8 ** m and cb.shape[0] == 12 * 4 ** m
        assert 2 * cp.shape[0] + cb.shape[0] == 4 * M.num_cells and np.all(cp != cm)
    shapes = [v for v in PUB["bsr_shapes"].values() if v["dim"] == 3]
    assert sorted((v["mesh"], v["p"]) for v in shapes) == [("S5", 3), ("S6", 2), ("S7", 1)]
    for v in shapes:
        m, nl = int(v["mesh"][1:]), _nloc3(v["p"])
        T = 6 * 8 ** m
        nnzb = T + (4 * T - 12 * 4 ** m)
        if m <= 5:
            (cp, _), _, _ = dg.UniformTetMesh(m).faces()
            assert nnzb == T + 2 * cp.shape[0]
        assert v["dofs_per_element"] == nl and v["dofs"] == T * nl
        assert v["bsr_nnz"] == nnzb and v["nnz"] == nnzb * nl * nl


def test_tet_mesh_hierarchy():
    """Each cell of S_3 lies in the S_k cell that parent_cells names (barycentric
    coordinates >= 0), every S_k cell holds 8^(3-k) of them; C_k cells hold the cells whose
    centroids they contain."""
    M = dg.UniformTetMesh(3)
    X = M.vertices[M.cells]
    for k in range(3):
        C = dg.UniformTetMesh(k)
        par = M.parent_cells(k)
        Y = C.vertices[C.cells[par]]
        Tm = np.stack([Y[:, 1] - Y[:, 0], Y[:, 2] - Y[:, 0], Y[:, 3] - Y[:, 0]], -1)
        lam = np.linalg.solve(Tm[:, None], (X - Y[:, :1])[..., None])[..., 0]
        bary = np.concatenate([1 - lam.sum(-1, keepdims=True), lam], -1)
        assert bary.min() > -1e-12
        assert np.all(np.bincount(par, minlength=C.num_cells) == 8 ** (3 - k))
        q = M.quad_cells(k)
        assert np.all(np.bincount(q, minlength=8 ** k) == 6 * 8 ** (3 - k))
    with pytest.raises(ValueError):
        dg.reference_nodes_tet(4)


@pytest.mark.parametrize("p,levels,order", [(1, (2, 3), 1.6), (2, (1, 2, 3), 2.9), (3, (1, 2), 3.5)])
def test_sipg_3d_symmetric_spd_and_converges(p, levels, order):
    """The 3D SIPG system is symmetric and positive definite, and its solution converges to
    the constant-coefficient 3D test case's exact solution at order p + 1 in L2 (the
    observed rates on S1-S3 approach it from below: p = 1 1.39, 1.76; p = 2 3.16, 3.25;
    p = 3 3.82)."""
    errs = []
    for m in levels:
        pr = dg.assemble_sipg(dg.UniformTetMesh(m), p, dg.constant_coefficient_3d)
        A = pr.A
        assert pr.block_size == _nloc3(p) and A.shape[0] == 6 * 8 ** m * _nloc3(p)
        assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
        if m == levels[0]:
            assert np.linalg.eigvalsh(A.toarray()).min() > 0
        u = spla.spsolve(A.tocsc(), pr.b)
        errs.append(dg.error_norm(pr, u, 2, dg.constant_coefficient_3d))
    rates = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert rates[-1] > order, rates


@pytest.mark.parametrize("kind", ["additive", "hybrid"])
def test_3d_experiment_with_the_oracle(kind):
    """The experiment set-up on a 3D system (S3, subdomains S1, coarse C0 / S0): the
    oracle's Schwarz CG converges, and its first residual is ||b||."""
    pr = dg.assemble_sipg(dg.UniformTetMesh(3), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S1", "S0")
    assert k == 64 * 4 and spc.sum() == 48 and spc.shape[0] == 6
    O = BlockSchwarzOracle(A, k, spc, kind=kind)
    x, md = cg(lambda v: A @ v, b, precond=O.apply, T0=O.T0 if kind == "hybrid" else None,
               rtol=1e-9, maxiter=500)
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b) * 1.01
    perm2, A2, b2, k2, spc2 = dg.dg_experiment(pr, "C1", "C0")
    assert spc2.tolist() == [8] and k2 == 3072 // 8 * 4
