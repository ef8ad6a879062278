"""Golden data for the differentiated solve at 48^3 (the differentiation workload).

theta ~ 0.5 N(0,1) per element (seeded stream [0, 1]; 0.5 is the standard deviation),
theta_obs from stream [1, 1], f ~ U(-1,1) per unknown (stream [0, 2]); u_obs = A(theta_obs)^-1 f.
Writes adjoint48.npz with, from FP64 direct solves (SuperLU) in oracle/adjoint.py:
  J          0.5 ||u - u_obs||^2 at theta
  grad       dJ/dtheta (48^3 values), adjoint method
  fd         central differences along 32 directions v_i ~ N(0,1) (rng seed 7), eps = 1e-6,
             each J(theta +- eps v) solved to the last digits (refinement with the factor of
             A(theta))
  u_obs      the observation (so the GPU test need not re-solve it)

    python tests/golden/make_adjoint_golden.py [N]
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dd-solvers_amd"))

from dd_solvers import problems as P  # noqa: E402
from oracle.adjoint import fd_directional, objective_and_gradient  # noqa: E402


def inputs(N: int, seed: int = 0):
    theta = 0.5 * np.random.default_rng([seed, 1]).standard_normal(N ** 3)
    theta_obs = 0.5 * np.random.default_rng([seed + 1, 1]).standard_normal(N ** 3)
    f = P.uniform_rhs((N - 1) ** 3, seed)
    return theta, theta_obs, f


def directions(N: int, count: int = 32):
    return np.random.default_rng(7).standard_normal((count, N ** 3))


def main(N: int = 48):
    t0 = time.time()
    theta, theta_obs, f = inputs(N)
    A_obs = P.assemble_q1_3d(N, np.exp(theta_obs).reshape(N, N, N)).tocsc()
    u_obs = spla.splu(A_obs, permc_spec="MMD_AT_PLUS_A").solve(f)
    print(f"u_obs {time.time() - t0:.0f}s", flush=True)
    A = P.assemble_q1_3d(N, np.exp(theta).reshape(N, N, N)).tocsc()
    lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A")
    J, g, u, lam = objective_and_gradient(N, theta, f, u_obs, factor=lu)
    print(f"J={J:.6e} |grad|={np.linalg.norm(g):.6e} {time.time() - t0:.0f}s", flush=True)
    dirs = directions(N)
    fd = fd_directional(N, theta, f, u_obs, dirs, eps=1e-6, precond_lu=lu)
    ad = dirs @ g
    print("max |fd - adjoint| / max |adjoint| =", float(np.abs(fd - ad).max() / np.abs(ad).max()))
    np.savez_compressed(os.path.join(HERE, f"adjoint{N}.npz"), J=np.array(J), grad=g, fd=fd,
                        u_obs=u_obs, eps=np.array(1e-6), N=np.array(N))
    print(f"done {time.time() - t0:.0f}s")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 48)
