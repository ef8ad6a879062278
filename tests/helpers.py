"""Shared fixtures-as-functions for the tests (inputs only; the oracle lives in oracle/)."""

import os

import numpy as np
import scipy.sparse as sps

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dg24():
    """The reference's 24-unknown SIPG system (src/dd_solvers/tests/test_solvers.py:25-108)."""
    d = np.load(os.path.join(GOLDEN, "dg24.npz"))
    A = sps.csr_matrix((d["values"], d["col_indices"].astype(np.int64), d["crow_indices"].astype(np.int64)),
                       shape=(24, 24))
    return A, d["load_vector"], d["fine_to_solvers"], d["solvers_to_coarse"], int(d["dofs_per_solver"])


def dg24_permuted():
    """DG-24 in the subdomain numbering of asm_permutation (restated with numpy)."""
    A, b, f2s, s2c, k = dg24()
    f2d = np.arange(24).reshape(8, 3)
    n_s = s2c.shape[0]
    solver_dofs = f2d[np.argsort(f2s, kind="stable")].reshape(n_s, -1)
    perm = solver_dofs[np.argsort(s2c, kind="stable")].reshape(-1)
    Ap = A[perm][:, perm].tocsr()
    Ap.sort_indices()
    spc = np.bincount(s2c, minlength=s2c.max() + 1)
    return A, b, perm, Ap, k, spc


def bsr8():
    d = np.load(os.path.join(GOLDEN, "bsr8.npz"))
    return d["crow_indices"], d["col_indices"], d["values"]
