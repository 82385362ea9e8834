"""Solves with JacobianLinearOperator on the GPU kernels.

The operator's products reach the Krylov kernels matrix-free (CG on a
PSD Hessian, GMRES and BiCGStab on a nonsymmetric Jacobian, jac = "fwd"
and "bwd") and the dense kernels through materialise (LU, QR least
squares); each answer matches numpy's float64 solve with the closed-form
Jacobian, to the solver's tolerance.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx

pytestmark = pytest.mark.gpu


def _jacobian_problem(device, n, seed, m=None):
    g = torch.Generator(device="cpu").manual_seed(seed)
    m = m or n
    a = torch.randn(m, dtype=torch.float64, generator=g)
    b = torch.randn(m, n, dtype=torch.float64, generator=g) / np.sqrt(n) + 3.0 * torch.eye(m, n, dtype=torch.float64)
    c = torch.randn(m, n, dtype=torch.float64, generator=g) * 0.1
    x = torch.randn(n, dtype=torch.float64, generator=g)
    J = b + c * (2 * x)[None, :]
    a, b, c, x = (t.to(device) for t in (a, b, c, x))
    fn = lambda x, _: a + b @ x + c @ (x ** 2)
    return fn, x, J.numpy()


@pytest.mark.parametrize("jac", ["fwd", "bwd"])
@pytest.mark.parametrize("solver", ["gmres", "bicgstab"])
def test_jacobian_krylov_solves(device, jac, solver):
    fn, x, J = _jacobian_problem(device, 40, seed=1)
    op = lx.JacobianLinearOperator(fn, x, jac=jac)
    rhs = np.random.default_rng(2).standard_normal(40)
    s = lx.GMRES(1e-10, 1e-10, restart=20) if solver == "gmres" else lx.BiCGStab(1e-10, 1e-10)
    sol = lx.linear_solve(op, torch.as_tensor(rhs, device=device), s)
    assert int(sol.result) == lx.RESULTS.successful
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(J, rhs), rtol=1e-8, atol=1e-9)


def test_jacobian_cg_on_hessian(device):
    # the Jacobian of the gradient of f(x) = 1/2 x^T M x + sum cosh(x) / 10
    # is M + diag(cosh(x)) / 10: symmetric positive definite
    n = 50
    rng = np.random.default_rng(3)
    q = rng.standard_normal((n, n))
    m = torch.as_tensor(q @ q.T / n + np.eye(n), device=device)
    grad = lambda x, _: m @ x + torch.sinh(x) / 10
    x0 = torch.as_tensor(rng.standard_normal(n), device=device)
    hess = (m + torch.diag(torch.cosh(x0)) / 10).cpu().numpy()
    op = lx.JacobianLinearOperator(grad, x0, tags=lx.positive_semidefinite_tag)
    rhs = rng.standard_normal(n)
    sol = lx.linear_solve(op, torch.as_tensor(rhs, device=device), lx.CG(1e-11, 1e-11))
    assert int(sol.result) == lx.RESULTS.successful
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(hess, rhs), rtol=1e-8, atol=1e-9)
    # the same system through the dense path (materialised, Cholesky kernels)
    sol = lx.linear_solve(op, torch.as_tensor(rhs, device=device), lx.Cholesky())
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(hess, rhs), rtol=1e-10, atol=1e-11)


def test_jacobian_dense_solves(device):
    fn, x, J = _jacobian_problem(device, 30, seed=4)
    op = lx.JacobianLinearOperator(fn, x)
    rhs = np.random.default_rng(5).standard_normal(30)
    sol = lx.linear_solve(op, torch.as_tensor(rhs, device=device), lx.LU())
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(J, rhs), rtol=1e-10, atol=1e-11)
    # a tall Jacobian: least squares through QR
    fn, x, J = _jacobian_problem(device, 12, seed=6, m=30)
    op = lx.JacobianLinearOperator(fn, x)
    sol = lx.linear_solve(op, torch.as_tensor(rhs, device=device), lx.QR())
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.lstsq(J, rhs, rcond=None)[0], rtol=1e-10,
                               atol=1e-11)
