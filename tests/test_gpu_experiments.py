"""The experiment harness end to end on the GPU: a frame in the reference's schema whose
iterations, residuals and times are those of the solver it ran."""

import numpy as np
import pytest
import torch

from oracle.pcg import cg as oracle_cg
from oracle.schwarz import BlockSchwarzOracle

import dd_solvers as dds
from dd_solvers import dg
from dd_solvers.experiments import COLUMNS, Experiment, ExperimentFactory, UniformMeshes

pytestmark = pytest.mark.gpu


def test_factory_runs_the_reference_protocol():
    print("experiments: p=1 S6, S4/S3 and S4/S4, Hybrid and Additive", flush=True)
    fam = UniformMeshes(2, 6)
    f = ExperimentFactory(dg.continuous_coefficient_2d, fam, polynomial_degree=1,
                          setup_repetitions=2, solution_warmup_steps=1,
                          solution_measurement_steps=2, solution_repetitions=3,
                          error_continuous_norms=(2,))
    f.add(Experiment("S6", "S4", "S3", dds.CG(dds.HybridSchwarz(
        torch.float64, dds.Inv(), dds.SparseCholesky()), rtol=1e-9, maxiter=2000)))
    f.add(Experiment("S6", "S4", "S4", dds.CG(dds.AdditiveSchwarz(
        torch.float64, dds.Inv(torch.float16), dds.DenseCholesky()), rtol=1e-9, maxiter=2000)))
    f.add(Experiment("S6", "S4", "S4", dds.CG(dds.HybridSchwarz(
        torch.float64, dds.Inv(), dds.DenseCholesky()), rtol=1e-9, maxiter=3)))  # raises
    df = f.run()
    cols = COLUMNS[:16] + ["error L2 norm"] + COLUMNS[16:]
    assert list(df.columns) == cols
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(6), 1)
    # the first row against the oracle on the same permuted system
    perm, A, b, k, spc = dg.dg_experiment(pr, "S4", "S3")
    O = BlockSchwarzOracle(A, k, spc, kind="hybrid")
    _, mo = oracle_cg(lambda v: A @ v, b, precond=O.apply, T0=O.T0, rtol=1e-9, maxiter=2000)
    r0 = df.iloc[0]
    assert r0["exception"] is None
    assert abs(r0["metadata"]["iterations"] - mo["iterations"]) <= 1
    assert r0["residual norm"] <= 1e-9 * np.linalg.norm(pr.b) * 1.01
    assert len(r0["solve times"]) == 3 and all(t > 0 for t in r0["solve times"])
    assert len(r0["solver setup times"]) == 2
    assert r0["solver"] == "CG(HybridSchwarz(torch.float64, Inv(), SparseCholesky), bsr_matmul=True)"
    assert r0["DoFs"] == pr.A.shape[0] and r0["matrix nnz"] == pr.A.nnz
    assert 0 < r0["error L2 norm"] < 0.01
    r1 = df.iloc[1]
    assert r1["exception"] is None and r1["residual norm"] <= 1e-9 * np.linalg.norm(pr.b) * 1.01
    # a run that does not converge records the exception and its metadata
    r2 = df.iloc[2]
    assert "converge" in r2["exception"].lower() or "iteration" in r2["exception"].lower()
    assert r2["metadata"]["iterations"] in (2, 3)
