"""Oracle restatement of `CG.solve` (src/dd_solvers/dd_solvers/solvers.py:192-303).

Pins (tests/test_oracle.py):
  * condition-number known answer diag(0.5, 4, 21), b = 1, rtol 1e-9 -> 42.0
    (src/dd_solvers/tests/test_solvers.py:179-184);
  * without a preconditioner, CG on SPD n x n terminates in <= n steps and matches
    scipy.sparse.linalg.spsolve; the recurrence residual equals b - A x_i;
  * with M = A^{-1} exactly one iteration is taken;
  * the raise quirk (i + 1 == maxiter raises even after a break at maxiter - 1).
"""

from __future__ import annotations

import numpy as np

__all__ = ["ConvergenceNotAchieved", "cg"]


class ConvergenceNotAchieved(Exception):
    """solvers.py:28-37."""

    def __init__(self, metadata):
        self.metadata = metadata
        super().__init__(f"Convergence not achieved in {metadata.get('iterations')} iterations.")


def cg(matvec, b, precond=None, T0=None, rtol=1e-7, atol=0.0, maxiter=1000,
       estimate_cond=False, x0=None, record_iterates=False, dot=None):
    """Preconditioned CG exactly as solvers.py:228-297.

    * atol_eff = max(atol, rtol * ||b||) (:229); maxiter None/0 -> 10 n (:231);
    * x = x0 (mutated in place, :232) else zeros; Hybrid adds T0(b) (:233-234);
    * r = b - A x (:235); the norm is logged and checked at the top of every iteration,
      before preconditioning (:244-247);
    * p0 = copy(z0); p = p * beta + z with beta = rho_i / rho_{i-1} (:256-262);
    * alpha = rho / (p . q); x += alpha p; r -= alpha q (:264-268);
    * iterations = len(norms) - 1 (:276); raise iff i + 1 == maxiter (:294-295).
    Returns (x, metadata); metadata gains "iterates" when record_iterates is set.
    dot (default numpy's): the inner product for rho, p . q and the residual norm, so that
    the same recurrence can be run with its reductions summed in another order.
    """
    b = np.asarray(b, dtype=np.float64)
    norm = np.linalg.norm if dot is None else (lambda v: float(np.sqrt(dot(v, v))))
    dot = dot or np.dot
    atol_eff = max(atol, rtol * float(np.linalg.norm(b)))
    n = b.shape[0]
    maxiter = maxiter or n * 10
    x = x0 if x0 is not None else np.zeros_like(b)
    if T0 is not None:
        x += T0(b)
    r = b - matvec(x)
    rho_prev = None
    p = None
    norms = []
    iterates = [x.copy()] if record_iterates else None
    alphas, betas = [], []
    i = -1
    for i in range(maxiter):
        res = float(norm(r))
        norms.append(res)
        if res < atol_eff:
            break
        z = precond(r) if precond is not None else r
        rho = float(np.dot(r, z))
        if i > 0:
            beta = rho / rho_prev
            p *= beta
            p += z
        else:
            p = np.array(z, dtype=np.float64, copy=True)
        q = matvec(p)
        alpha = rho / float(np.dot(p, q))
        x += alpha * p
        r -= alpha * q
        rho_prev = rho
        if record_iterates:
            iterates.append(x.copy())
        if estimate_cond:
            if i > 0:
                betas.append(beta)
            alphas.append(alpha)
    meta = {"iterations": len(norms) - 1, "residual norms": norms,
            "preconditioner metadata": [{} for _ in range(len(norms) - 1)]}
    if record_iterates:
        meta["iterates"] = iterates
    if estimate_cond:
        m = len(alphas)
        T = np.zeros((m, m))
        for j in range(m):
            T[j, j] = 1.0 / alphas[j]
            if j > 0:
                T[j, j] += betas[j - 1] / alphas[j - 1]
                T[j - 1, j] = T[j, j - 1] = betas[j - 1] ** 0.5 / alphas[j - 1]
        s = np.linalg.svd(T, compute_uv=False)
        meta["condition number estimate"] = float(s[0] / s[-1])
    if i + 1 == maxiter:
        raise ConvergenceNotAchieved(meta)
    return x, meta
