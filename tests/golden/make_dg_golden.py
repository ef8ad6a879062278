"""Oracle residual histories of a full-size DG run for tests/test_gpu_dg.py.

Continuous-coefficient problem, p = 1, fine mesh S10 (6,291,456 DoFs), subdomains S8
(131,072 of 48 DoFs), coarse cells C6 (4,096; a coarse space the dense coarse solver
holds), FP64, rtol 1e-9, Additive and Hybrid Schwarz. The oracle (BlockSchwarzOracle +
oracle.pcg.cg) is the one that reproduces the reference's published S10 / S8 / S8 runs
(tests/test_dg.py). Writes tests/golden/dg_s10_s8_c6.json.

Run from the repository root:  python tests/golden/make_dg_golden.py
"""

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dd-solvers_amd"))

from oracle.pcg import cg  # noqa: E402
from oracle.schwarz import BlockSchwarzOracle  # noqa: E402
from dd_solvers import dg  # noqa: E402


def main():
    pr = dg.assemble_sipg(dg.UniformTriangleMesh(10), 1)
    perm, A, b, k, spc = dg.dg_experiment(pr, "S8", "C6")
    M = BlockSchwarzOracle(A, k, spc)
    out = {"problem": "continuous coefficient 2D, p=1, S10 / S8 / C6, FP64, rtol 1e-9",
           "dofs": int(A.shape[0]), "dofs_per_solver": int(k), "n_coarse": int(spc.shape[0]),
           "b_norm": float(np.linalg.norm(b))}
    for kind in ("additive", "hybrid"):
        M.kind = kind
        x, md = cg(lambda v: A @ v, b, precond=M.apply, T0=M.T0 if kind == "hybrid" else None,
                   rtol=1e-9, maxiter=2000)
        out[kind] = {"iterations": md["iterations"], "residual_norms": md["residual norms"],
                     "x_norm": float(np.linalg.norm(x)),
                     "x_sample": x[:: max(1, x.shape[0] // 64)].tolist()}
        print(kind, md["iterations"], flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dg_s10_s8_c6.json")
    with open(path, "w") as fh:
        json.dump(out, fh)
    print("wrote", path)


if __name__ == "__main__":
    main()
