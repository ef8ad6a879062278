"""The experiment harness (dd_solvers.experiments) on the host: the result schema of the
reference's results/*.csv, the mesh-family maps, and the continuous error norm."""

import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from dd_solvers import dg
from dd_solvers.experiments import COLUMNS, UniformMeshes

PUB = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dg_published.json")))


def test_columns_are_the_reference_csv_header():
    assert COLUMNS == PUB["columns"]


def test_mesh_family_maps():
    fam = UniformMeshes(2, 6)
    assert fam.name == "uniform(2D,6)"
    m6 = dg.UniformTriangleMesh(6)
    assert np.array_equal(fam.get_mapping("S6", "S4"), m6.parent_cells(4))
    assert np.array_equal(fam.get_mapping("S6", "C3"), m6.quad_cells(3))
    # S_k -> S_j composes, C_k -> C_j follows the 2^k grid
    assert np.array_equal(fam.get_mapping("S4", "S2")[fam.get_mapping("S6", "S4")],
                          fam.get_mapping("S6", "S2"))
    assert np.array_equal(fam.get_mapping("C3", "C1")[fam.get_mapping("S6", "C3")],
                          fam.get_mapping("S6", "C1"))
    with pytest.raises(NotImplementedError):
        UniformMeshes(3, 4)


def test_error_norm_converges_at_order_p_plus_1():
    """L2 error of the SIPG solution against sin(4 pi x) sin(4 pi y): order p + 1."""
    for p, levels in ((1, (5, 6, 7)), (2, (3, 4, 5))):
        errs = []
        for m in levels:
            pr = dg.assemble_sipg(dg.UniformTriangleMesh(m), p)
            errs.append(dg.error_norm(pr, spla.spsolve(pr.A.tocsc(), pr.b), 2))
        rate = np.log2(errs[-2] / errs[-1])
        assert rate > p + 1 - 0.25, (p, errs)


def test_error_norm_of_the_interpolant_is_small():
    """u_h = nodal values of the exact solution: the L1 and L2 errors are interpolation
    errors, O(h^2) for p = 1, and zero error gives zero."""
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(6), 1)
    xy = pr.node_coords.reshape(-1, 2)
    u = dg.continuous_coefficient_2d.solution(xy[:, 0], xy[:, 1])
    e2 = dg.error_norm(pr, u, 2)
    assert 0 < e2 < 5e-3
    assert dg.error_norm(pr, u, 1) <= e2 * 1.0001  # |Omega| = 1: L1 <= L2


def test_bench_runs_every_published_family():
    """bench.py's reference runs name exactly the published families: the precision
    experiment's 12 (default and on request) and the parallelism experiment's 30, each with
    the published solvers."""
    import json
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    pub = json.load(open(os.path.join(root, "tests", "golden", "dg_published.json")))["solves"]
    prec = {(r["p"], r["fine"], r["solvers"], r["coarse"]) for r in pub if r["experiment"] == "precisions"}
    ours = {(p, f"S{fine}", sm, cm) for p, fine, sm, cm in bench.REF_EXPERIMENTS + bench.REF_EXPERIMENTS_LARGE}
    assert ours == prec
    par = {(r["solvers"], r["coarse"]) for r in pub if r["experiment"] == "parallelism"}
    assert set(bench.REF_PARALLELISM) == par and len(bench.REF_PARALLELISM) == 30
    par_solvers = {(r["kind"], r["precision"], r["inv"]) for r in pub if r["experiment"] == "parallelism"}
    assert par_solvers == set(bench.PAR_SOLVERS)
    prec_solvers = {(r["kind"], r["precision"], r["inv"]) for r in pub if r["experiment"] == "precisions"}
    assert set(bench.REF_SOLVERS) <= prec_solvers


def test_bench_all_solver_configurations_are_the_published_ones():
    import json
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    pub = json.load(open(os.path.join(root, "tests", "golden", "dg_published.json")))["solves"]
    prec_solvers = {(r["kind"], r["precision"], r["inv"]) for r in pub if r["experiment"] == "precisions"}
    assert set(bench.REF_SOLVERS_ALL) == prec_solvers and len(bench.REF_SOLVERS_ALL) == 14
