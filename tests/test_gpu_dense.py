"""GPU parity of the LU and Cholesky kernels against the oracle.

Tolerances: float64 solves must match the float64 oracle to 1e-10 relative
(matrices with condition number < 1000 or ~50 for SPD); float32 solves must
have normwise backward error below 64 * n * eps32 against the exact float64
data, and identical pivot sequences to the oracle's getrf on tie-free input.
"""

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd import _kernels as K
from lineax_amd import workloads as W
from oracle import lineax_oracle as O
from tests.helpers import poisson_1d, well_conditioned

pytestmark = pytest.mark.gpu

EPS32 = np.finfo(np.float32).eps


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 9, 16, 17, 31, 32, 33, 48, 64, 65, 100, 130])
def test_lu_factor_matches_oracle_f64(device, rng, n):
    batch = 3
    mats = [well_conditioned(rng, n) for _ in range(batch)]
    lu, piv, info = K.lu_factor(_t(np.stack(mats), device))
    torch.cuda.synchronize()
    for i in range(batch):
        lu_ref, piv_ref, info_ref = O.lu_factor(mats[i])
        np.testing.assert_array_equal(piv[i].cpu().numpy(), piv_ref)
        np.testing.assert_allclose(lu[i].cpu().numpy(), lu_ref, rtol=1e-10, atol=1e-11)
        assert int(info[i]) == info_ref == 0


@pytest.mark.parametrize("n", [1, 5, 32, 64, 65, 129])
@pytest.mark.parametrize("trans", [0, 1])
def test_lu_solve_f64(device, rng, n, trans):
    a = well_conditioned(rng, n)
    b = rng.standard_normal((4, n))
    lu, piv, _ = K.lu_factor(_t(a, device))
    x = K.lu_solve(lu, piv, _t(b, device), trans).cpu().numpy()
    for i in range(4):
        ref = np.linalg.solve(a.T if trans else a, b[i])
        np.testing.assert_allclose(x[i], ref, rtol=1e-10, atol=1e-10)


def test_small_lu_config_batch_f32(device):
    # The 1024 x (32 x 32) configuration: pivots equal the oracle's and the
    # float32 solutions are backward stable against the exact data.
    a, b = W.small_lu_batch()
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device, torch.float32)), _t(b, device, torch.float32),
                          lx.LU())
    x = sol.value.cpu().numpy().astype(np.float64)
    assert sol.value.shape == (1024, 32) and sol.value.dtype == torch.float32
    assert torch.all(sol.result == 0)
    _, piv, _ = K.lu_factor(_t(a, device, torch.float32))
    piv = piv.cpu().numpy()
    for i in range(0, 1024, 37):
        _, piv_ref, _ = O.lu_factor(a[i])
        np.testing.assert_array_equal(piv[i], piv_ref)
    etas = [O.normwise_backward_error(a[i], x[i], b[i]) for i in range(1024)]
    assert max(etas) < 64 * 32 * EPS32


def test_lu_state_reuse_and_broadcast(device, rng):
    a = well_conditioned(rng, 20)
    op = lx.MatrixLinearOperator(_t(a, device))
    state = lx.LU().init(op, {})
    bs = rng.standard_normal((5, 20))
    sol = lx.linear_solve(op, _t(bs, device), lx.LU(), state=state)
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(a, bs.T).T, rtol=1e-10, atol=1e-10)


def test_lu_transpose_state(device, rng):
    a = well_conditioned(rng, 12)
    b = rng.standard_normal(12)
    op = lx.MatrixLinearOperator(_t(a, device))
    st, opts = lx.LU().transpose(lx.LU().init(op, {}), {})
    sol = lx.linear_solve(op.T, _t(b, device), lx.LU(), state=st, options=opts)
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(a.T, b), rtol=1e-10, atol=1e-10)


def test_lu_singular_and_nonfinite_results(device, rng):
    a = rng.standard_normal((2, 6, 6))
    a[0, 2, :] = 0.0  # singular system 0
    b = rng.standard_normal((2, 6))
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.LU(), throw=False)
    assert sol.result.tolist() == [lx.RESULTS.singular, lx.RESULTS.successful]
    b2 = b.copy()
    b2[0, 1] = np.nan
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b2, device), lx.LU(), throw=False)
    assert sol.result.tolist() == [lx.RESULTS.nonfinite_input, lx.RESULTS.successful]
    with pytest.raises(RuntimeError, match="non-finite"):
        lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.LU())


def test_reference_kat_2x2_gpu(device):
    a = _t([[1.0, 5.0], [-2.0, -2.0]], device)
    x = lx.linear_solve(lx.MatrixLinearOperator(a), _t([3.0, 4.0], device)).value
    np.testing.assert_allclose(x.cpu().numpy(), [-3.25, 1.25], rtol=1e-14)


@pytest.mark.parametrize("n", [1, 4, 8, 13, 32, 40, 64, 65, 100, 181, 200, 256, 300])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cholesky_matches_oracle(device, rng, n, dtype):
    batch = 2
    mats = [well_conditioned(rng, n, kind="psd") for _ in range(batch)]
    a = np.stack(mats)
    if dtype == torch.float32:
        a = a.astype(np.float32).astype(np.float64)
    b = rng.standard_normal((batch, n))
    u, info = K.chol_factor(_t(a, device, dtype))
    x = K.chol_solve(u, _t(b, device, dtype)).cpu().numpy()
    assert torch.all(info == 0)
    for i in range(batch):
        u_ref, _ = O.cho_factor(a[i])
        ref = O.cho_solve(u_ref, b[i])
        if dtype == torch.float64:
            np.testing.assert_allclose(u[i].cpu().numpy(), u_ref, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(x[i], ref, rtol=1e-10, atol=1e-10)
        else:
            assert O.normwise_backward_error(a[i], x[i], b[i]) < 64 * n * EPS32
            du = np.abs(np.triu(u[i].cpu().numpy()) - u_ref).max() / np.abs(u_ref).max()
            assert du < 16 * n * EPS32
        lower = np.tril(u[i].cpu().numpy(), -1)
        assert np.all(lower == 0)


@pytest.mark.parametrize("n", [6, 70, 200, 300])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cholesky_failure_is_singular(device, rng, n, dtype):
    # float32 orders 65..256 take the register-tiled factor, the rest the
    # lane-per-row, LDS and global-memory tiers
    a = well_conditioned(rng, n, kind="psd")
    a[n // 2, n // 2] = -5.0
    op = lx.MatrixLinearOperator(_t(a, device, dtype), lx.positive_semidefinite_tag)
    sol = lx.linear_solve(op, _t(rng.standard_normal(n), device, dtype), lx.Cholesky(), throw=False)
    assert int(sol.result) == lx.RESULTS.singular
    u, info = K.chol_factor(_t(a, device, dtype))
    assert int(info) > 0 and torch.all(torch.isnan(u))


@pytest.mark.parametrize("n", [5, 50, 120])
def test_cholesky_nsd(device, rng, n):
    a = well_conditioned(rng, n, kind="nsd")
    b = rng.standard_normal(n)
    op32 = lx.MatrixLinearOperator(_t(a, device, torch.float32), lx.negative_semidefinite_tag)
    x32 = lx.linear_solve(op32, _t(b, device, torch.float32), lx.Cholesky()).value.double().cpu().numpy()
    a32 = a.astype(np.float32).astype(np.float64)
    assert O.normwise_backward_error(a32, x32, b.astype(np.float32).astype(np.float64)) < 64 * n * EPS32
    op = lx.MatrixLinearOperator(_t(a, device), lx.negative_semidefinite_tag)
    sol = lx.linear_solve(op, _t(b, device), lx.Cholesky())
    np.testing.assert_allclose(sol.value.cpu().numpy(), np.linalg.solve(a, b), rtol=1e-10, atol=1e-10)
    auto = lx.linear_solve(op, _t(b, device))
    np.testing.assert_allclose(auto.value.cpu().numpy(), np.linalg.solve(a, b), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("n,dtype", [(40, torch.float64), (150, torch.float32), (240, torch.float32)])
def test_cholesky_reads_upper_triangle_only(device, rng, n, dtype):
    a = well_conditioned(rng, n, kind="psd")
    junk = a.copy()
    junk[np.tril_indices(n, -1)] = 1e30
    u1, _ = K.chol_factor(_t(a, device, dtype))
    u2, _ = K.chol_factor(_t(junk, device, dtype))
    assert torch.equal(u1, u2)


def test_gemv_matches(device, rng):
    # (5000, 300) and (4099, 37) split the transposed product over row chunks
    # (the 16-byte path and the scalar path); the rest run in one chunk
    for m, n in [(1, 1), (5, 7), (64, 64), (300, 129), (5000, 300), (4099, 37)]:
        a = rng.standard_normal((3, m, n))
        x = rng.standard_normal((3, n))
        y = K.gemv(_t(a, device), _t(x, device)).cpu().numpy()
        np.testing.assert_allclose(y, np.einsum("bij,bj->bi", a, x), rtol=1e-12, atol=1e-12)
        xt = rng.standard_normal((3, m))
        yt = K.gemv(_t(a, device), _t(xt, device), trans=True).cpu().numpy()
        np.testing.assert_allclose(yt, np.einsum("bij,bi->bj", a, xt), rtol=1e-12, atol=1e-12)
        # transposed views run the other product on the same storage, no copy
        ta = _t(a, device)
        yv = K.gemv(ta.mT, _t(xt, device)).cpu().numpy()
        np.testing.assert_allclose(yv, np.einsum("bij,bi->bj", a, xt), rtol=1e-12, atol=1e-12)
        yvt = K.gemv(ta.mT, _t(x, device), trans=True).cpu().numpy()
        np.testing.assert_allclose(yvt, np.einsum("bij,bj->bi", a, x), rtol=1e-12, atol=1e-12)
        op_t = lx.MatrixLinearOperator(ta[0]).transpose()
        np.testing.assert_allclose(op_t.mv(_t(xt[0], device)).cpu().numpy(), a[0].T @ xt[0], rtol=1e-12,
                                   atol=1e-12)


# --- differentiation: lineax's JVP (x' = A^{-1}(b' - A' x)) and transpose rules.


@pytest.mark.parametrize("solver", ["lu", "cholesky"])
def test_grad_closed_forms(device, rng, solver):
    n = 9
    kind = "psd" if solver == "cholesky" else "general"
    a = well_conditioned(rng, n, kind=kind)
    b = rng.standard_normal(n)
    g = rng.standard_normal(n)
    tags = lx.positive_semidefinite_tag if solver == "cholesky" else ()
    A = _t(a, device).requires_grad_(True)
    B = _t(b, device).requires_grad_(True)
    s = lx.LU() if solver == "lu" else lx.Cholesky()
    x = lx.linear_solve(lx.MatrixLinearOperator(A, tags), B, s).value
    (x * _t(g, device)).sum().backward()
    xs = np.linalg.solve(a, b)
    lam = np.linalg.solve(a.T, g)
    np.testing.assert_allclose(B.grad.cpu().numpy(), lam, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(A.grad.cpu().numpy(), -np.outer(lam, xs), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(x.detach().cpu().numpy(), xs, rtol=1e-10, atol=1e-10)


def test_jvp_closed_form(device, rng):
    import torch.autograd.forward_ad as fwAD

    n = 7
    a = well_conditioned(rng, n)
    b = rng.standard_normal(n)
    da = rng.standard_normal((n, n))
    db = rng.standard_normal(n)
    with fwAD.dual_level():
        A = fwAD.make_dual(_t(a, device), _t(da, device))
        B = fwAD.make_dual(_t(b, device), _t(db, device))
        x = lx.linear_solve(lx.MatrixLinearOperator(A), B, lx.LU()).value
        primal, tangent = fwAD.unpack_dual(x)
    xs = np.linalg.solve(a, b)
    np.testing.assert_allclose(primal.cpu().numpy(), xs, rtol=1e-10)
    np.testing.assert_allclose(tangent.cpu().numpy(), np.linalg.solve(a, db - da @ xs), rtol=1e-9, atol=1e-10)


def test_invert_operator(device, rng):
    a = well_conditioned(rng, 10, kind="psd")
    inv = lx.invert(lx.MatrixLinearOperator(_t(a, device), lx.positive_semidefinite_tag))
    assert lx.is_positive_semidefinite(inv)
    v = rng.standard_normal(10)
    np.testing.assert_allclose(inv.mv(_t(v, device)).cpu().numpy(), np.linalg.solve(a, v), rtol=1e-10, atol=1e-10)


def test_poisson_nsd_with_cholesky(device):
    a = poisson_1d(100)
    b = np.linspace(-1, 1, 100)
    op = lx.MatrixLinearOperator(_t(a, device), (lx.negative_semidefinite_tag, lx.symmetric_tag))
    x = lx.linear_solve(op, _t(b, device), lx.Cholesky()).value
    np.testing.assert_allclose(x.cpu().numpy(), np.linalg.solve(a, b), rtol=1e-9, atol=1e-9)


# --- the fused small-LU solve (lxa_lu_small_solve): one launch, bit-equal to
# lxa_lu_factor + lxa_lu_solve, and the lazy LU state built on it


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("n", [1, 5, 8, 13, 16, 31, 32, 33, 64])
def test_lu_small_fused_bit_equal(device, rng, dtype, n):
    batch = 37
    a = rng.standard_normal((batch, n, n))
    a[3] = 0.0  # singular: zero pivots, info, non-finite x like the pair
    a[4, :, 0] = 0.0
    a[5] = np.round(a[5])  # ties in the pivot search
    b = rng.standard_normal((batch, n))
    A, B = _t(a, device).to(dtype), _t(b, device).to(dtype)
    x, lu, piv, info = K.lu_small_solve(A, B, want_factors=True)
    lu_ref, piv_ref, info_ref = K.lu_factor(A)
    x_ref = K.lu_solve(lu_ref, piv_ref, B, trans=0)
    assert torch.equal(piv, piv_ref) and torch.equal(info, info_ref)
    assert torch.equal(torch.nan_to_num(lu, nan=7.0), torch.nan_to_num(lu_ref, nan=7.0))
    assert torch.equal(torch.nan_to_num(x, nan=7.0, posinf=8.0, neginf=9.0),
                       torch.nan_to_num(x_ref, nan=7.0, posinf=8.0, neginf=9.0))
    x2, lu2, piv2, _ = K.lu_small_solve(A, B)
    assert lu2 is None and piv2 is None and torch.equal(torch.nan_to_num(x2, nan=7.0, posinf=8.0, neginf=9.0),
                                                         torch.nan_to_num(x, nan=7.0, posinf=8.0, neginf=9.0))
    # operators one element off 16-byte alignment take the row-per-lane loads
    base = torch.empty(batch * n * n + 1, dtype=dtype, device=device)
    A_off = base[1:].view(batch, n, n)
    A_off.copy_(A)
    x3, lu3, piv3, _ = K.lu_small_solve(A_off, B, want_factors=True)
    assert torch.equal(piv3, piv) and torch.equal(torch.nan_to_num(lu3, nan=7.0), torch.nan_to_num(lu, nan=7.0))
    assert torch.equal(torch.nan_to_num(x3, nan=7.0, posinf=8.0, neginf=9.0),
                       torch.nan_to_num(x, nan=7.0, posinf=8.0, neginf=9.0))
    # and the solves are backward stable (random Gaussian operators: judged by
    # the normwise backward error, not the forward error)
    good = [i for i in range(batch) if i not in (3, 4)]
    xs = x.double().cpu().numpy()
    eps = float(torch.finfo(dtype).eps)
    for i in good[:8]:
        ai = A[i].double().cpu().numpy()
        bi = B[i].double().cpu().numpy()
        assert O.normwise_backward_error(ai, xs[i], bi) <= 64 * n * eps


def test_lu_lazy_state_reuse_and_transpose(device, rng):
    n, batch = 24, 6
    a = np.stack([well_conditioned(rng, n) for _ in range(batch)])
    b1, b2 = rng.standard_normal((batch, n)), rng.standard_normal((batch, n))
    op = lx.MatrixLinearOperator(_t(a, device))
    solver = lx.LU()
    state = solver.init(op, {})
    x1, _, _ = solver.compute(state, _t(b1, device), {})  # fused: factors formed here
    x2, _, _ = solver.compute(state, _t(b2, device), {})  # reuses them
    for i in range(batch):
        np.testing.assert_allclose(x1[i].cpu().numpy(), np.linalg.solve(a[i], b1[i]), rtol=1e-10, atol=1e-11)
        np.testing.assert_allclose(x2[i].cpu().numpy(), np.linalg.solve(a[i], b2[i]), rtol=1e-10, atol=1e-11)
    # a transposed state before any solve forms the factors itself
    state = solver.init(op, {})
    st_t, _ = solver.transpose(state, {})
    xt, _, _ = solver.compute(st_t, _t(b1, device), {})
    for i in range(batch):
        np.testing.assert_allclose(xt[i].cpu().numpy(), np.linalg.solve(a[i].T, b1[i]), rtol=1e-10, atol=1e-11)
    # the state returned by linear_solve is reusable
    sol = lx.linear_solve(op, _t(b1, device), solver)
    again = lx.linear_solve(op, _t(b2, device), solver, state=sol.state)
    np.testing.assert_allclose(again.value.cpu().numpy(), x2.cpu().numpy(), rtol=0, atol=0)
