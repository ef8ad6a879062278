"""Oracle for the differentiated solve: J(theta) = 0.5 ||u(theta) - u_obs||^2 with
A(theta) u = f, A the Q1 stiffness of -div(exp(theta) grad u) on an N^3 grid.

The reference has no autograd (it solves with fixed matrices). This restates the adjoint
method: A is symmetric, so the adjoint solve A lam = u - u_obs uses the same operator, and
dJ/dtheta_e = -lam_e^T (dA/dtheta_e) u_e = -k_e h lam_e^T Khat u_e, where Khat is the
unit-coefficient Q1 element stiffness of the unit cube and h = 1/N (3D scaling h^(d-2)).

Pins (tests/test_oracle.py): the element formula against dA/dtheta_e assembled by brute
force on a tiny grid, and the whole gradient against central finite differences with
direct solves (eps = 1e-6) on 32 random directions.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla


def _problems():
    import dd_solvers.problems as P  # input generation, no solver code
    return P


def element_gradient(N: int, kappa: np.ndarray, lam: np.ndarray, u: np.ndarray) -> np.ndarray:
    """-k_e h lam_e^T Khat u_e for every element (natural unknown order, Dirichlet = 0)."""
    P = _problems()
    K = P.q1_reference_stiffness()
    nodes = P.q1_element_nodes(N)
    lz = np.concatenate([lam, [0.0]])[np.where(nodes >= 0, nodes, lam.shape[0])]
    uz = np.concatenate([u, [0.0]])[np.where(nodes >= 0, nodes, u.shape[0])]
    q = np.einsum("ea,ab,eb->e", lz, K, uz)
    return -kappa.reshape(-1) * q / N


def objective_and_gradient(N: int, theta: np.ndarray, f: np.ndarray, u_obs: np.ndarray,
                           factor=None):
    """(J, dJ/dtheta, u, lam) by a direct FP64 solve (SuperLU) of A(theta)."""
    P = _problems()
    kappa = np.exp(theta).reshape(N, N, N)
    A = P.assemble_q1_3d(N, kappa).tocsc()
    lu = factor if factor is not None else spla.splu(A, permc_spec="MMD_AT_PLUS_A")
    u = lu.solve(f)
    r = u - u_obs
    J = 0.5 * float(r @ r)
    lam = lu.solve(r)
    return J, element_gradient(N, kappa, lam, u), u, lam


def objective(N: int, theta: np.ndarray, f: np.ndarray, u_obs: np.ndarray, precond_lu=None,
              tol: float = 1e-15) -> float:
    """J(theta). With `precond_lu` (a factorization of a nearby A) the solve is CG
    preconditioned by that factorization, iterated to the last digits; else a fresh
    direct factorization."""
    P = _problems()
    A = P.assemble_q1_3d(N, np.exp(theta).reshape(N, N, N)).tocsr()
    if precond_lu is None:
        u = spla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A").solve(f)
    else:
        u = precond_lu.solve(f)
        for _ in range(20):  # iterative refinement with the nearby factorization
            r = f - A @ u
            if np.linalg.norm(r) <= tol * np.linalg.norm(f):
                break
            u = u + precond_lu.solve(r)
    r = u - u_obs
    return 0.5 * float(r @ r)


def fd_directional(N: int, theta, f, u_obs, directions, eps: float = 1e-6, precond_lu=None):
    """Central differences (J(theta + eps v) - J(theta - eps v)) / (2 eps) for each v."""
    out = []
    for v in directions:
        jp = objective(N, theta + eps * v, f, u_obs, precond_lu)
        jm = objective(N, theta - eps * v, f, u_obs, precond_lu)
        out.append((jp - jm) / (2 * eps))
    return np.array(out)
