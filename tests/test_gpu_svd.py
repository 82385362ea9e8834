"""GPU parity of the one-sided Jacobi SVD and the SVD solver.

The singular values match numpy's LAPACK SVD to 1e-12 relative (float64)
or 1e-5 (float32); U and V have orthonormal columns and U diag(s) V^T
reproduces A to 1e-12 (float64); the solver equals the oracle's
pseudoinverse solve (tall, wide, square, rank-deficient, batched, the
reference's non-square KATs) and reports the numerical rank; the transposed
state solves with A^T.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd import _kernels as K
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (3, 5), (16, 16), (33, 17), (17, 33), (100, 64), (64, 129),
                                   (200, 200)])
def test_svd_factors_float64(device, rng, shape):
    a = rng.standard_normal(shape)
    u, s, vt, sweeps = K.svd(_t(a, device))
    u, s, vt = (t.cpu().numpy() for t in (u, s, vt))
    s_ref = np.linalg.svd(a, compute_uv=False)
    np.testing.assert_allclose(s, s_ref, rtol=1e-12, atol=1e-12 * s_ref[0])
    assert np.all(np.diff(s) <= 0)
    k = min(shape)
    np.testing.assert_allclose(u.T @ u, np.eye(k), atol=1e-12)
    np.testing.assert_allclose(vt @ vt.T, np.eye(k), atol=1e-12)
    np.testing.assert_allclose(u @ np.diag(s) @ vt, a, atol=1e-12 * np.abs(a).max() * max(shape))
    assert 0 < int(sweeps.item()) < K.SVD_MAX_SWEEPS


def test_svd_factors_float32_and_batched(device, rng):
    a = rng.standard_normal((3, 40, 25)).astype(np.float32)
    u, s, vt, _ = K.svd(_t(a, device, torch.float32))
    for i in range(3):
        s_ref = np.linalg.svd(a[i].astype(np.float64), compute_uv=False)
        np.testing.assert_allclose(s[i].cpu().numpy(), s_ref, rtol=1e-5, atol=1e-5 * s_ref[0])
        rec = u[i].double().cpu().numpy() @ np.diag(s[i].double().cpu().numpy()) @ vt[i].double().cpu().numpy()
        np.testing.assert_allclose(rec, a[i], atol=1e-5 * np.abs(a[i]).max() * 40)


@pytest.mark.parametrize("shape", [(30, 12), (12, 30), (20, 20)])
def test_svd_solver_full_rank(device, rng, shape):
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.SVD())
    x_ref, rank = O.svd_solve(a, b)
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-10, atol=1e-11)
    assert int(sol.stats["rank"]) == rank == min(shape)


def test_svd_solver_rank_deficient_and_zero(device, rng):
    a = rng.standard_normal((15, 4)) @ rng.standard_normal((4, 10))
    b = rng.standard_normal(15)
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.SVD())
    x_ref, rank = O.svd_solve(a, b)
    assert int(sol.stats["rank"]) == rank == 4
    np.testing.assert_allclose(sol.value.cpu().numpy(), x_ref, rtol=1e-8, atol=1e-10)
    z = lx.linear_solve(lx.MatrixLinearOperator(torch.zeros((3, 4), dtype=torch.float64, device=device)),
                        torch.ones(3, dtype=torch.float64, device=device), lx.SVD())
    assert int(z.stats["rank"]) == 0 and bool((z.value == 0).all())


def test_svd_solver_kats_batch_and_transpose(device, rng):
    # lineax tests/test_singular.py:100-145
    a1 = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    a2 = np.array([[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]])
    for a, y in ((a1, [3.0, 4.0]), (a2, [3.0, 4.0, 5.0])):
        for solver in (lx.SVD(), lx.AutoLinearSolver(well_posed=False)):
            x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(y, device), solver).value
            np.testing.assert_allclose(x.cpu().numpy(), np.linalg.lstsq(a, y, rcond=None)[0], rtol=1e-12,
                                       atol=1e-12)
    a = rng.standard_normal((4, 9, 6))
    b = rng.standard_normal((4, 9))
    x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.SVD()).value.cpu().numpy()
    for i in range(4):
        np.testing.assert_allclose(x[i], O.svd_solve(a[i], b[i])[0], rtol=1e-10, atol=1e-11)
    a = rng.standard_normal((14, 9))
    solver = lx.SVD()
    state = solver.init(lx.MatrixLinearOperator(_t(a, device)), {})
    st_t, _ = solver.transpose(state, {})
    c = rng.standard_normal(9)
    x_t, _, _ = solver.compute(st_t, _t(c, device), {})
    np.testing.assert_allclose(x_t.cpu().numpy(), np.linalg.lstsq(a.T, c, rcond=None)[0], rtol=1e-10, atol=1e-11)
