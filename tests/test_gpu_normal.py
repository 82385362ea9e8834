"""GPU parity of Normal (lineax/_solver/normal.py) against the oracle.

Normal(Cholesky) in float64 matches the oracle's normal-equations solve and
numpy's lstsq to 1e-9 relative (tall: least squares; wide: minimum norm;
condition numbers below ~30, squared to ~1000 by the normal equations, as
lineax's tests/helpers.py:76-78 cut them).  Normal(CG) matches lstsq to its
tolerance.  The transposed state solves with A^T; the derivative through a
square full-rank Normal solve is the solve's closed form.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd import _kernels as K
from oracle import lineax_oracle as O
from tests.helpers import poisson_1d

pytestmark = pytest.mark.gpu


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def _matrix(rng, m, n, cond_max=30.0):
    while True:
        a = rng.standard_normal((m, n))
        if np.linalg.cond(a) < cond_max:
            return a


@pytest.mark.parametrize("shape", [(6, 3), (3, 6), (40, 40), (200, 37), (37, 200), (1, 5), (5, 1)])
def test_normal_cholesky_vs_oracle(device, rng, shape):
    m, n = shape
    a = _matrix(rng, m, n)
    b = rng.standard_normal(m)
    x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.Normal(lx.Cholesky())).value
    x = x.cpu().numpy()
    np.testing.assert_allclose(x, O.normal_solve(a, b), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-9, atol=1e-10)


def test_normal_nonsquare_kats(device):
    # lineax tests/test_singular.py:100-145 (answers: lstsq of the same data)
    a1 = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    y1 = np.array([3.0, 4.0])
    a2 = np.array([[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]])
    y2 = np.array([3.0, 4.0, 5.0])
    for a, y in ((a1, y1), (a2, y2)):
        x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(y, device), lx.Normal(lx.Cholesky())).value
        np.testing.assert_allclose(x.cpu().numpy(), np.linalg.lstsq(a, y, rcond=None)[0], rtol=1e-11, atol=1e-12)


def test_normal_float32(device, rng):
    a = _matrix(rng, 300, 60)
    b = rng.standard_normal(300)
    x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device, torch.float32)), _t(b, device, torch.float32),
                        lx.Normal(lx.Cholesky())).value
    assert x.dtype == torch.float32
    x = x.double().cpu().numpy()
    # the normal-equations residual to float32 accuracy (kappa(A)^2 ~ 1e3)
    assert O.lstsq_backward_error(a, x, b) <= 1e-5


@pytest.mark.parametrize("shape", [(50, 20), (20, 50)])
def test_normal_cg_vs_lstsq(device, rng, shape):
    m, n = shape
    a = _matrix(rng, m, n, cond_max=10.0)
    b = rng.standard_normal(m)
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device),
                          lx.Normal(lx.CG(rtol=1e-10, atol=1e-10)))
    assert int(sol.result) == lx.RESULTS.successful
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-7, atol=1e-8)


def test_normal_cg_max_steps_only(device, rng):
    # lineax tests/test_solve.py:174-194: runs with max_steps only
    op = lx.MatrixLinearOperator(_t(poisson_1d(100), device), (lx.negative_semidefinite_tag, lx.symmetric_tag))
    s = lx.linear_solve(op, _t(rng.standard_normal(100), device), lx.Normal(lx.CG(0.0, 0.0, max_steps=2)),
                        throw=False)
    assert int(s.stats["num_steps"]) == 2 and int(s.result) == lx.RESULTS.successful


def test_normal_transpose_state(device, rng):
    for m, n in ((40, 25), (25, 40)):
        a = _matrix(rng, m, n)
        op = lx.MatrixLinearOperator(_t(a, device))
        solver = lx.Normal(lx.Cholesky())
        state = solver.init(op, {})
        st_t, _ = solver.transpose(state, {})
        c = rng.standard_normal(n)
        x_t, _, _ = solver.compute(st_t, _t(c, device), {})
        np.testing.assert_allclose(x_t.cpu().numpy(), np.linalg.lstsq(a.T, c, rcond=None)[0], rtol=1e-9,
                                   atol=1e-10)


def test_normal_grad_square_closed_form(device, rng):
    a = _matrix(rng, 10, 10) + 3 * np.eye(10)
    b = rng.standard_normal(10)
    g = rng.standard_normal(10)
    A = _t(a, device).requires_grad_(True)
    B = _t(b, device).requires_grad_(True)
    x = lx.linear_solve(lx.MatrixLinearOperator(A), B, lx.Normal(lx.Cholesky())).value
    (x * _t(g, device)).sum().backward()
    lam = np.linalg.solve(a.T, g)
    xs = np.linalg.solve(a, b)
    np.testing.assert_allclose(x.detach().cpu().numpy(), xs, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(B.grad.cpu().numpy(), lam, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(A.grad.cpu().numpy(), -np.outer(lam, xs), rtol=1e-8, atol=1e-9)


def test_normal_preconditioned_cg_exact_inverse(device, rng):
    # lineax tests/test_adjoint.py:80-130: an exact inverse as preconditioner
    # converges in a few steps, and the Jacobian of the solve is A^-1
    n = 10
    a = rng.uniform(size=(n, n)) + np.diag(np.arange(n) ** 6.0)
    b = rng.uniform(size=n)
    op = lx.MatrixLinearOperator(_t(a, device))
    m = lx.MatrixLinearOperator(_t(np.linalg.inv(a), device))
    sol = lx.linear_solve(op, _t(b, device), lx.Normal(lx.CG(1e-10, 1e-10, max_steps=4)),
                          options={"preconditioner": m})
    assert int(sol.result) == lx.RESULTS.successful
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(a, b), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("shape_a,shape_b", [((37, 200), (200, 37)), ((1, 5), (5, 1)), ((130, 129), (129, 257)),
                                             ((4, 64, 33), (4, 33, 70)), ((3, 20, 9), (9, 20)),
                                             ((2, 1, 8, 8), (5, 8, 8))])
def test_batched_gemm_vs_float64(device, rng, shape_a, shape_b):
    # the Gram product runs on lxa_gemm_batched: float32 inputs, float32
    # accumulation; each entry within 4 k eps of the float64 product scaled
    # by sum |a||b|
    a = rng.standard_normal(shape_a)
    b = rng.standard_normal(shape_b)
    c = K.gemm(_t(a, device, torch.float32), _t(b, device, torch.float32)).double().cpu().numpy()
    a32 = a.astype(np.float32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    exact = np.matmul(a32, b32)
    bound = 4 * shape_a[-1] * np.finfo(np.float32).eps * np.matmul(np.abs(a32), np.abs(b32))
    assert c.shape == exact.shape
    assert np.all(np.abs(c - exact) <= bound + 1e-30)


@pytest.mark.parametrize("ta,tb", [(True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape_a,shape_b", [((37, 200), (200, 37)), ((130, 129), (129, 257)),
                                             ((4, 64, 33), (4, 33, 70)), ((3, 20, 9), (9, 20)), ((1, 3), (3, 1))])
def test_batched_gemm_transposed_operands(device, rng, shape_a, shape_b, ta, tb):
    # transposed views (.mT of contiguous storage) go to the kernel as
    # LXA_GEMM_TRANS_A / _B and are read in place; same bound as above
    a = rng.standard_normal(shape_a)
    b = rng.standard_normal(shape_b)
    A = _t(np.swapaxes(a, -1, -2).copy(), device, torch.float32).mT if ta else _t(a, device, torch.float32)
    B = _t(np.swapaxes(b, -1, -2).copy(), device, torch.float32).mT if tb else _t(b, device, torch.float32)
    assert A.shape == a.shape and B.shape == b.shape
    c = K.gemm(A, B).double().cpu().numpy()
    a32 = a.astype(np.float32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    exact = np.matmul(a32, b32)
    bound = 4 * shape_a[-1] * np.finfo(np.float32).eps * np.matmul(np.abs(a32), np.abs(b32))
    assert c.shape == exact.shape
    assert np.all(np.abs(c - exact) <= bound + 1e-30)


def test_normal_float32_gram_on_gemm_batched(device, rng):
    # tall float32 problems: the Gram matrix comes from the matrix-core GEMM,
    # the solve from the Cholesky kernels
    a = np.stack([_matrix(rng, 120, 40) for _ in range(3)])
    b = rng.standard_normal((3, 120))
    A = _t(a, device, torch.float32)
    gram = lx.MatrixLinearOperator(A[0]).transpose() @ lx.MatrixLinearOperator(A[0])
    g = gram.as_matrix().double().cpu().numpy()
    a0 = a[0].astype(np.float32).astype(np.float64)
    bound = 4 * 120 * np.finfo(np.float32).eps * (np.abs(a0).T @ np.abs(a0))
    assert np.all(np.abs(g - a0.T @ a0) <= bound)
    for i in range(3):
        x = lx.linear_solve(lx.MatrixLinearOperator(A[i]), _t(b[i], device, torch.float32),
                            lx.Normal(lx.Cholesky())).value
        assert O.lstsq_backward_error(a[i], x.double().cpu().numpy(), b[i]) <= 1e-5
