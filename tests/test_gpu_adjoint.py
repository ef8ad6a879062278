"""GPU checks of the differentiated solve (adjoint PCG + element-gradient kernel).

Bars: the element kernel equals the oracle's formula to rounding; the autograd gradient
agrees with the FP64 CPU direct-solve gradient and with central finite differences
(eps = 1e-6, 32 random directions) to 1e-6 relative.
"""

import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from helpers import GOLDEN
from oracle.adjoint import element_gradient, objective_and_gradient

from dd_solvers import problems as P
from dd_solvers.adjoint import Q1DiffSolve

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _inputs(N, seed=0):
    theta = 0.5 * np.random.default_rng([seed, 1]).standard_normal(N ** 3)
    theta_obs = 0.5 * np.random.default_rng([seed + 1, 1]).standard_normal(N ** 3)
    f = P.uniform_rhs((N - 1) ** 3, seed)
    return theta, theta_obs, f


def test_element_gradient_kernel_matches_oracle():
    N = 10
    op = Q1DiffSolve(N, box=3, overlap=1)
    rng = np.random.default_rng(3)
    theta = 0.5 * rng.standard_normal(N ** 3)
    lam = rng.standard_normal((N - 1) ** 3)
    u = rng.standard_normal((N - 1) ** 3)
    perm = op.decomposition.perm
    g = op.element_gradient(torch.as_tensor(theta, device=dev), torch.as_tensor(lam[perm], device=dev),
                            torch.as_tensor(u[perm], device=dev))
    ref = element_gradient(N, np.exp(theta).reshape(N, N, N), lam, u)
    assert np.allclose(g.cpu().numpy(), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_assembled_matrix_matches_host_assembly():
    N = 12
    op = Q1DiffSolve(N, box=4, overlap=1)
    theta, _, _ = _inputs(N)
    A = op.matrix(torch.as_tensor(theta, device=dev)).to_dense().cpu().numpy()
    A0 = P.assemble_q1_3d(N, np.exp(theta).reshape(N, N, N)).toarray()
    perm = op.decomposition.perm
    assert np.allclose(A, A0[perm][:, perm], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("storage", [torch.float64, torch.float32])
def test_autograd_matches_direct_solve_gradient(storage):
    N = 16
    theta, theta_obs, f = _inputs(N)
    A_obs = P.assemble_q1_3d(N, np.exp(theta_obs).reshape(N, N, N)).tocsc()
    u_obs = spla.spsolve(A_obs, f)
    J_ref, g_ref, u_ref, lam_ref = objective_and_gradient(N, theta, f, u_obs)
    op = Q1DiffSolve(N, box=4, overlap=1, storage=storage, rtol=1e-13)
    th = torch.as_tensor(theta, device=dev).requires_grad_(True)
    ft = torch.as_tensor(f, device=dev).requires_grad_(True)
    u = op(th, ft)
    J = 0.5 * ((u - torch.as_tensor(u_obs, device=dev)) ** 2).sum()
    J.backward()
    assert np.isclose(J.item(), J_ref, rtol=1e-10)
    g = th.grad.cpu().numpy()
    assert np.linalg.norm(g - g_ref) <= 1e-8 * np.linalg.norm(g_ref)
    assert np.linalg.norm(ft.grad.cpu().numpy() - lam_ref) <= 1e-8 * np.linalg.norm(lam_ref)


def test_adjoint_48_against_direct_gradient_and_finite_differences():
    """The differentiation workload: 48^3, theta ~ 0.5 N(0,1), FP32 local inverses, FP64 CG
    to rtol 1e-12, against the FP64 CPU direct-solve gradient and central differences
    (eps = 1e-6, 32 directions) stored by tests/golden/make_adjoint_golden.py."""
    path = os.path.join(GOLDEN, "adjoint48.npz")
    if not os.path.exists(path):
        pytest.skip("golden data not generated")
    gd = np.load(path)
    N = int(gd["N"])
    theta, _, f = _inputs(N)
    op = Q1DiffSolve(N, box=8, overlap=1, storage=torch.float32, rtol=1e-12)
    th = torch.as_tensor(theta, device=dev).requires_grad_(True)
    u = op(th, torch.as_tensor(f, device=dev))
    J = 0.5 * ((u - torch.as_tensor(gd["u_obs"], device=dev)) ** 2).sum()
    J.backward()
    g = th.grad.cpu().numpy()
    assert np.isclose(J.item(), float(gd["J"]), rtol=1e-9)
    rel = np.linalg.norm(g - gd["grad"]) / np.linalg.norm(gd["grad"])
    assert rel < 1e-6, rel
    dirs = np.random.default_rng(7).standard_normal((32, N ** 3))
    ad = dirs @ g
    fd = gd["fd"]
    assert np.all(np.abs(ad - fd) <= 1e-6 * np.abs(fd).max()), np.abs(ad - fd).max() / np.abs(fd).max()


def test_two_forwards_then_one_backward():
    """Two solves at different theta before one backward (the gradient of J1 + J2): each
    node's adjoint runs with the operator of its own forward, checked against the sum of
    the two FP64 direct-solve gradients."""
    N = 12
    theta1, theta_obs, f = _inputs(N, seed=0)
    theta2, _, _ = _inputs(N, seed=5)
    u_obs = spla.spsolve(P.assemble_q1_3d(N, np.exp(theta_obs).reshape(N, N, N)).tocsc(), f)
    _, g1, _, _ = objective_and_gradient(N, theta1, f, u_obs)
    _, g2, _, _ = objective_and_gradient(N, theta2, f, u_obs)
    op = Q1DiffSolve(N, box=4, overlap=1, storage=torch.float64, rtol=1e-13)
    t1 = torch.as_tensor(theta1, device=dev).requires_grad_(True)
    t2 = torch.as_tensor(theta2, device=dev).requires_grad_(True)
    uo = torch.as_tensor(u_obs, device=dev)
    ft = torch.as_tensor(f, device=dev)
    u1 = op(t1, ft)
    u2 = op(t2, ft)  # replaces op.solver before either backward runs
    J = 0.5 * ((u1 - uo) ** 2).sum() + 0.5 * ((u2 - uo) ** 2).sum()
    J.backward()
    assert np.linalg.norm(t1.grad.cpu().numpy() - g1) <= 1e-8 * np.linalg.norm(g1)
    assert np.linalg.norm(t2.grad.cpu().numpy() - g2) <= 1e-8 * np.linalg.norm(g2)
