"""Experiment harness with the reference's result schema, on the DG systems of dd_solvers.dg.

The reference runs its experiments through `ExperimentFactory` (src/experiments/experiments/
experiments.py:34-318) on DOLFINx discretisations and writes one pandas row per experiment;
its results/*.csv files are those rows. This module keeps the class names, constructor
arguments, the run protocol and the column names, so a run here produces a frame that
lines up with the reference's CSVs row for row:

  * the matrix and DoF map are copied to the GPU, the DoFs are renumbered with
    `asm_permutation` and the matrix permuted with `permute_csr_tensor` (:194-219);
  * `solver.setup(matrix, dofs_per_solver=, solvers_per_coarse=, block_size=)` (:221-235);
  * the right-hand side is the load vector or U(-1, 1) values (:237-245), permuted;
  * the last of `setup_repetitions` set-ups measures `solution_repetitions` rounds of
    `solution_warmup_steps` untimed and `solution_measurement_steps` timed solves with
    events (:247-266); every phase is timed with events as in the reference (:190-275);
  * the residual norm ||A x - b|| is computed on the host in FP64 (:281-283);
  * a failing run records the exception text and its metadata (:108-139).

What differs: the discretisation is dd_solvers.dg (the reference's S_m meshes, any p,
no DOLFINx), so `mesh_family` is `UniformMeshes(2, m)` built from its meshes, and error
norms come from dg.error_norm (exact solution at quadrature points) instead of a DOLFINx
form against its DG_{p+2} interpolant.
"""

from __future__ import annotations

import copy
import datetime
import gc
from typing import NamedTuple

import numpy as np
import torch

from . import dg
from .csr import asm_permutation, csr_to_torch, inverse_permutation, permute_csr_tensor
from .solvers import SparseSolver

__all__ = ["Experiment", "ExperimentFactory", "UniformMeshes", "COLUMNS"]

# the column order of the reference's results/*.csv (error-norm columns, when requested,
# follow "residual norm")
COLUMNS = ["datetime", "test case", "p", "mesh family", "fine m", "solvers m", "coarse m",
           "solver", "random rhs", "solution warmup steps", "solution measurement steps",
           "solution repetitions", "DoFs", "matrix nnz", "metadata", "residual norm",
           "setup metadata", "solve times", "matrix and DOF map copy times",
           "matrix permute times", "solver setup times", "rhs copy and permute times",
           "solution permute and copy times", "exception"]


class UniformMeshes:
    """The reference's `UniformMeshes(d, m)` family (meshing.py:197-283) for d = 2: the S_k
    meshes (k Plaza refinements of the unit square) and the C_k quadrilateral meshes, with
    the cell maps between them."""

    def __init__(self, d: int, m: int):
        if d != 2:
            raise NotImplementedError("only the 2D family is built (the reference's DG runs)")
        self.d, self.m = d, m
        self._meshes = {}

    @property
    def name(self) -> str:
        return f"uniform({self.d}D,{self.m})"

    def __getitem__(self, idx: str) -> dg.UniformTriangleMesh:
        if idx[0] != "S":
            raise ValueError(f"only S meshes carry a discretisation here (got {idx})")
        k = int(idx[1:])
        if k not in self._meshes:
            self._meshes[k] = dg.UniformTriangleMesh(k)
        return self._meshes[k]

    def get_mapping(self, from_mesh: str, to_mesh: str) -> np.ndarray:
        """Cell of `to_mesh` containing each cell of `from_mesh` (meshing.py:251-263)."""
        if from_mesh[0] == "S":
            return self[from_mesh].level_map(to_mesh)
        if from_mesh[0] == "C" and to_mesh[0] == "C":
            k, j = int(from_mesh[1:]), int(to_mesh[1:])
            q = 2 ** k
            J, I = np.divmod(np.arange(q * q), q)
            return (J >> (k - j)) * (2 ** j) + (I >> (k - j))
        raise ValueError(f"Unsupported mapping from {from_mesh} to {to_mesh}")


class Experiment(NamedTuple):
    fine_m: str
    solvers_m: str | None
    coarse_m: str | None
    solver: SparseSolver


class ExperimentFactory:
    """Runs `Experiment`s and returns a pandas DataFrame in the reference's schema."""

    def __init__(self, test_case: dg.DGTestCase, mesh_family: UniformMeshes,
                 polynomial_degree: int = 1, setup_repetitions: int = 3,
                 solution_warmup_steps: int = 0, solution_measurement_steps: int = 1,
                 solution_repetitions: int = 10, problem_precision=np.float64,
                 error_continuous_norms: tuple = (), random_rhs: bool = False,
                 device: str | torch.device = "cuda"):
        if setup_repetitions < 1 or solution_warmup_steps < 0 or solution_measurement_steps < 1 \
                or solution_repetitions < 1:
            raise AssertionError("bad repetition counts")
        self.test_case = test_case
        self.mesh_family = mesh_family
        self.polynomial_degree = polynomial_degree
        self.setup_repetitions = setup_repetitions
        self.solution_warmup_steps = solution_warmup_steps
        self.solution_measurement_steps = solution_measurement_steps
        self.solution_repetitions = solution_repetitions
        self.problem_precision = problem_precision
        self.error_continuous_norms = tuple(error_continuous_norms)
        self.random_rhs = random_rhs
        self.device = torch.device(device)
        self.experiments: list[Experiment] = []

    def add(self, experiment: Experiment) -> None:
        self.experiments.append(experiment)

    def run(self):
        import pandas as pd

        problems = {}
        for fine_m in sorted({e.fine_m for e in self.experiments}):
            pr = dg.assemble_sipg(self.mesh_family[fine_m], self.polynomial_degree, self.test_case)
            if self.problem_precision != np.float64:
                pr.A = pr.A.astype(self.problem_precision)
                pr.b = pr.b.astype(self.problem_precision)
            problems[fine_m] = pr
        rows = [self._repeat(e, problems[e.fine_m]) for e in self.experiments]
        return pd.DataFrame(rows)

    def _repeat(self, experiment: Experiment, problem: dg.DGProblem) -> dict:
        """experiments.py:108-178: set-up repetitions, times gathered into lists."""
        results, exception, exception_metadata = [], None, None
        for rep in range(self.setup_repetitions):
            try:
                results.append(self._run(experiment, problem, rep == self.setup_repetitions - 1))
            except KeyboardInterrupt:
                raise
            except Exception as e:  # recorded in the row, as the reference does
                exception = str(e)
                exception_metadata = getattr(e, "metadata", None)
            finally:
                gc.collect()
                torch.cuda.empty_cache()
            if exception is not None:
                break
        combined = {}
        if results:
            last = results[-1]
            combined = {k: v for k, v in last.items() if not k.endswith("time")}
            for key in [k for k in last if k.endswith("time")]:
                combined[f"{key}s"] = [r[key] for r in results if r.get(key) is not None]
            for key in [k for k in last if k.endswith("times")]:
                combined[key] = sum([r[key] for r in results if r.get(key) is not None], [])
        return {
            "datetime": datetime.datetime.now(tz=datetime.timezone.utc).isoformat(),
            "test case": self.test_case.name,
            "p": self.polynomial_degree,
            "mesh family": self.mesh_family.name,
            "fine m": experiment.fine_m,
            "solvers m": experiment.solvers_m,
            "coarse m": experiment.coarse_m,
            "solver": str(experiment.solver),
            "random rhs": self.random_rhs,
            "solution warmup steps": self.solution_warmup_steps,
            "solution measurement steps": self.solution_measurement_steps,
            "solution repetitions": self.solution_repetitions,
            "DoFs": problem.A.shape[0],
            "matrix nnz": problem.A.nnz,
            "metadata": exception_metadata,
            **combined,
            "exception": exception,
        }

    def _run(self, experiment: Experiment, problem: dg.DGProblem, measure: bool) -> dict:
        """experiments.py:180-318 on one set-up."""
        dev = self.device
        dd = experiment.coarse_m is not None and experiment.solvers_m is not None
        if dd:
            f2s = self.mesh_family.get_mapping(experiment.fine_m, experiment.solvers_m)
            s2c = self.mesh_family.get_mapping(experiment.solvers_m, experiment.coarse_m)
        solver = copy.deepcopy(experiment.solver)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(7)]

        ev[0].record()
        matrix = csr_to_torch(problem.A).to(dev)
        if dd:
            fine_to_solvers = torch.as_tensor(f2s).to(dev)
            solvers_to_coarse = torch.as_tensor(s2c).to(dev)
            fine_to_dofs = torch.as_tensor(problem.fine_to_dofs).to(dev)
        ev[1].record()
        if dd:
            perm = asm_permutation(fine_to_solvers, solvers_to_coarse, fine_to_dofs)
            inv_perm = inverse_permutation(perm)
            matrix = permute_csr_tensor(matrix, inv_perm)
        ev[2].record()
        kwargs = {}
        if dd:
            n_coarse = int(solvers_to_coarse.max().item()) + 1
            kwargs = dict(dofs_per_solver=int(torch.sum(fine_to_solvers == 0).item())
                          * fine_to_dofs.shape[1],
                          solvers_per_coarse=solvers_to_coarse.bincount(minlength=n_coarse))
        setup_metadata = solver.setup(matrix, **kwargs, block_size=problem.block_size)
        del matrix
        ev[3].record()
        if self.random_rhs:
            cpu_rhs = (np.random.rand(problem.b.shape[0]) - 0.5) * 2
        else:
            cpu_rhs = problem.b
        rhs = torch.as_tensor(cpu_rhs).to(dev)
        if dd:
            rhs = rhs[perm]
        ev[4].record()
        solve_times = []
        if measure:
            for _ in range(self.solution_repetitions):
                start = torch.cuda.Event(enable_timing=True)
                end = torch.cuda.Event(enable_timing=True)
                for _ in range(self.solution_warmup_steps):
                    solver.solve(rhs)
                start.record()
                for _ in range(self.solution_measurement_steps):
                    sol, metadata = solver.solve(rhs)
                end.record()
                torch.cuda.synchronize()
                solve_times.append(start.elapsed_time(end) / self.solution_measurement_steps)
        else:
            sol, metadata = solver.solve(rhs)
        ev[5].record()
        if dd:
            sol = sol[inv_perm]
        cpu_sol = sol.cpu().numpy()
        ev[6].record()
        torch.cuda.synchronize()
        solver.destroy()
        residual_norm = float(np.linalg.norm(problem.A @ cpu_sol - cpu_rhs))
        errors = {f"error L{q} norm": dg.error_norm(problem, cpu_sol, q, self.test_case)
                  for q in self.error_continuous_norms}
        return {
            "residual norm": residual_norm,
            **errors,
            "metadata": metadata,
            "setup metadata": setup_metadata,
            "matrix and DOF map copy time": ev[0].elapsed_time(ev[1]),
            "matrix permute time": ev[1].elapsed_time(ev[2]),
            "solver setup time": ev[2].elapsed_time(ev[3]),
            "rhs copy and permute time": ev[3].elapsed_time(ev[4]),
            "solve times": solve_times,
            "solution permute and copy time": ev[5].elapsed_time(ev[6]),
        }
