"""GPU parity of the Householder QR, QR solver and Triangular solver.

The factor is compared with the oracle's geqrf (same LAPACK reflector
convention, so R, the reflectors and tau agree entrywise) at float64 to
1e-11 relative to the column scale, and at float32 against the float64
factor of the same data to a float32 tolerance.  Solutions are compared
with the oracle's QR solve and with numpy's lstsq / minimum-norm solution.
At the largest sizes the checks are size-independent: the least-squares
optimality condition |A^T (A x - b)| and agreement with LAPACK (scipy).
"""

import numpy as np
import pytest
import scipy.linalg as sla
import torch

import lineax_amd as lx
from lineax_amd import _kernels as K
from oracle import lineax_oracle as O

pytestmark = pytest.mark.gpu


def _t(x, device, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def _raw_to_np(a, m, n):
    return a.cpu().numpy().astype(np.float64).reshape(m, n)


@pytest.mark.parametrize("shape", [(1, 1), (5, 3), (16, 16), (17, 9), (40, 17), (300, 100), (257, 257), (600, 33),
                                   (9, 4100), (4500, 20)])
def test_geqrf_f64_matches_oracle(device, rng, shape):
    m, n = shape
    a = rng.standard_normal(shape)
    ad, tau, _ = K.qr_factor(_t(a, device))
    h_ref = a.copy()
    tau_ref = np.zeros(min(m, n))
    O.lib().lxo_geqrf(m, n, O._dp(h_ref), n, O._dp(tau_ref))
    got = _raw_to_np(ad, m, n)
    scale = np.maximum(np.abs(h_ref).max(axis=0), 1.0)
    np.testing.assert_allclose(got / scale, h_ref / scale, rtol=0, atol=1e-11)
    np.testing.assert_allclose(tau.cpu().numpy(), tau_ref, rtol=0, atol=1e-12)


def test_geqrf_f64_matches_lapack_raw(device, rng):
    a = rng.standard_normal((70, 30))
    ad, tau, _ = K.qr_factor(_t(a, device))
    (h_ref, tau_ref), _ = sla.qr(a, mode="raw")
    np.testing.assert_allclose(_raw_to_np(ad, 70, 30), h_ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(tau.cpu().numpy(), tau_ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("shape", [(64, 16), (333, 200), (1000, 1000)])
def test_geqrf_f32_close_to_f64(device, rng, shape):
    m, n = shape
    a32 = rng.standard_normal(shape).astype(np.float32)
    ad, tau, _ = K.qr_factor(_t(a32, device, torch.float32))
    assert ad.dtype == torch.float32
    got = _raw_to_np(ad, m, n)
    (h_ref, tau_ref), _ = sla.qr(a32.astype(np.float64), mode="raw")
    r_got, r_ref = np.triu(got)[: min(m, n)], np.triu(h_ref)[: min(m, n)]
    tol = 2e-4 * np.sqrt(min(m, n))
    assert np.max(np.abs(r_got - r_ref)) <= tol * np.max(np.abs(r_ref))


def test_geqrf_zero_column_gives_zero_tau(device):
    a = np.zeros((6, 3))
    a[:, 1] = np.arange(6) + 1.0
    a[0, 2] = 2.0
    ad, tau, _ = K.qr_factor(_t(a, device))
    h_ref, tau_ref, _ = O.qr_factor(a)
    assert tau.cpu().numpy()[0] == 0.0
    np.testing.assert_allclose(_raw_to_np(ad, 6, 3), h_ref, atol=1e-14)
    np.testing.assert_allclose(tau.cpu().numpy(), tau_ref, atol=1e-15)


@pytest.mark.parametrize("shape", [(7, 7), (50, 20), (20, 50), (129, 128), (3, 1), (1, 3)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_qr_solve_vs_oracle(device, rng, shape, dtype):
    m, n = shape
    a = rng.standard_normal(shape)
    b = rng.standard_normal(m)
    if dtype == torch.float32:
        a = a.astype(np.float32).astype(np.float64)
        b = b.astype(np.float32).astype(np.float64)
    sol = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device, dtype)), _t(b, device, dtype), lx.QR())
    assert int(sol.result) == 0
    x_ref, res = O.solve_dense("qr", a, b)
    x_np = np.linalg.lstsq(a, b, rcond=None)[0]
    tol = 1e-10 if dtype == torch.float64 else 2e-3
    x = sol.value.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(x - x_ref)) <= tol * max(1.0, np.max(np.abs(x_ref)))
    assert np.max(np.abs(x - x_np)) <= tol * max(1.0, np.max(np.abs(x_np)))


def test_qr_nonsquare_kats(device):
    # lineax tests/test_singular.py:113-145 (answers are numpy lstsq of the same data)
    a1 = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    y1 = np.array([3.0, 4.0])
    a2 = np.array([[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]])
    y2 = np.array([3.0, 4.0, 5.0])
    for a, y in ((a1, y1), (a2, y2)):
        for solver in (lx.QR(), lx.AutoLinearSolver(well_posed=None)):
            x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(y, device), solver).value
            np.testing.assert_allclose(x.cpu().numpy(), np.linalg.lstsq(a, y, rcond=None)[0], rtol=1e-12,
                                       atol=1e-12)


def test_qr_batched_and_broadcast(device, rng):
    a = rng.standard_normal((3, 50, 20))
    b = rng.standard_normal((3, 50))
    x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device)), _t(b, device), lx.QR()).value.cpu().numpy()
    for i in range(3):
        np.testing.assert_allclose(x[i], np.linalg.lstsq(a[i], b[i], rcond=None)[0], rtol=1e-10, atol=1e-10)
    a1 = rng.standard_normal((30, 40))
    b1 = rng.standard_normal((4, 30))
    x1 = lx.linear_solve(lx.MatrixLinearOperator(_t(a1, device)), _t(b1, device), lx.QR()).value.cpu().numpy()
    for i in range(4):
        np.testing.assert_allclose(x1[i], np.linalg.lstsq(a1, b1[i], rcond=None)[0], rtol=1e-10, atol=1e-10)


def test_qr_transpose_state_reuse(device, rng):
    a = rng.standard_normal((40, 25))
    op = lx.MatrixLinearOperator(_t(a, device))
    solver = lx.QR()
    state = solver.init(op, {})
    st_t, _ = solver.transpose(state, {})
    c = rng.standard_normal(25)
    x_t, _, _ = solver.compute(st_t, _t(c, device), {})
    # A^T x = c, minimum norm
    np.testing.assert_allclose(x_t.cpu().numpy(), np.linalg.lstsq(a.T, c, rcond=None)[0], rtol=1e-10, atol=1e-10)


def test_qr_grad_square_closed_form(device, rng):
    a = rng.standard_normal((12, 12)) + 4 * np.eye(12)
    b = rng.standard_normal(12)
    g = rng.standard_normal(12)
    A = _t(a, device).requires_grad_(True)
    B = _t(b, device).requires_grad_(True)
    x = lx.linear_solve(lx.MatrixLinearOperator(A), B, lx.QR()).value
    (x * _t(g, device)).sum().backward()
    gb = np.linalg.solve(a.T, g)
    xs = np.linalg.solve(a, b)
    np.testing.assert_allclose(B.grad.cpu().numpy(), gb, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(A.grad.cpu().numpy(), -np.outer(gb, xs), rtol=1e-9, atol=1e-10)


def test_qr_large_least_squares_optimality(device, rng):
    m, n = 4096, 2048
    a = rng.standard_normal((m, n)).astype(np.float32)
    b = rng.standard_normal(m).astype(np.float32)
    x = lx.linear_solve(lx.MatrixLinearOperator(_t(a, device, torch.float32)), _t(b, device, torch.float32),
                        lx.QR()).value.cpu().numpy().astype(np.float64)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    x_ref = np.linalg.lstsq(a64, b64, rcond=None)[0]
    r = a64 @ x - b64
    # optimality: A^T r ~ 0 relative to |A| |r|
    assert np.linalg.norm(a64.T @ r) <= 1e-3 * np.linalg.norm(a64) * np.linalg.norm(r)
    assert np.linalg.norm(x - x_ref) <= 1e-3 * np.linalg.norm(x_ref)


def test_qr_uneven_batch_matches_per_system(device, rng):
    shapes = [(64, 16), (300, 299), (17, 17), (1000, 250), (5, 2), (129, 64)]
    nb = K.qr_block(torch.float64)
    descs = np.zeros(len(shapes), dtype=K.QR_DESC)
    a_off = t_off = w_off = tau_off = 0
    mats = []
    for i, (m, n) in enumerate(shapes):
        k = min(m, n)
        descs[i] = (a_off, tau_off, t_off, w_off, 0, m, n, n, 0)
        mats.append(rng.standard_normal((m, n)))
        a_off += m * n
        tau_off += k
        t_off += -(-k // nb) * nb * nb
        w_off += K.qr_work_elems(torch.float64, m, n)
    a = torch.cat([_t(x, device).reshape(-1) for x in mats])
    tau = torch.zeros(tau_off, dtype=torch.float64, device=device)
    tblk = torch.zeros(t_off, dtype=torch.float64, device=device)
    work = torch.zeros(w_off, dtype=torch.float64, device=device)
    dd = K.descs_to_device(descs, device)
    K.qr_factor_raw(dd, len(shapes), max(s[0] for s in shapes), max(s[1] for s in shapes),
                    max(min(s) for s in shapes), a, tau, tblk, work)
    got = a.cpu().numpy()
    gt = tau.cpu().numpy()
    for i, (m, n) in enumerate(shapes):
        h_ref = mats[i].copy()
        tau_ref = np.zeros(min(m, n))
        O.lib().lxo_geqrf(m, n, O._dp(h_ref), n, O._dp(tau_ref))
        off = int(descs[i]["a_offset"])
        hi = got[off: off + m * n].reshape(m, n)
        scale = np.maximum(np.abs(h_ref).max(axis=0), 1.0)
        np.testing.assert_allclose(hi / scale, h_ref / scale, rtol=0, atol=1e-11)
        to = int(descs[i]["tau_offset"])
        np.testing.assert_allclose(gt[to: to + min(m, n)], tau_ref, atol=1e-12)


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 200, 1031])
@pytest.mark.parametrize("lower", [False, True])
@pytest.mark.parametrize("trans", [False, True])
def test_trsv_matches_scipy(device, rng, n, lower, trans):
    a = rng.standard_normal((n, n)) / np.sqrt(n) + 2 * np.eye(n)
    a = np.tril(a) if lower else np.triu(a)
    b = rng.standard_normal((2, n))
    x = K.tri_solve(_t(a, device), _t(b, device), lower=lower, trans=trans, unit_diag=False).cpu().numpy()
    for i in range(2):
        ref = sla.solve_triangular(a, b[i], lower=lower, trans=1 if trans else 0)
        np.testing.assert_allclose(x[i], ref, rtol=1e-11, atol=1e-11)


def test_triangular_solver_tags_unit_and_transpose(device, rng):
    n = 70
    a = np.tril(rng.standard_normal((n, n)) / np.sqrt(n)) + 3 * np.eye(n)
    np.fill_diagonal(a, 1.0)
    b = rng.standard_normal(n)
    op = lx.MatrixLinearOperator(_t(a, device), (lx.lower_triangular_tag, lx.unit_diagonal_tag))
    x = lx.linear_solve(op, _t(b, device), lx.Triangular()).value.cpu().numpy()
    np.testing.assert_allclose(x, sla.solve_triangular(a, b, lower=True, unit_diagonal=True), rtol=1e-12,
                               atol=1e-12)
    solver = lx.Triangular()
    st, _ = solver.transpose(solver.init(op, {}), {})
    xt, _, _ = solver.compute(st, _t(b, device), {})
    np.testing.assert_allclose(xt.cpu().numpy(), sla.solve_triangular(a, b, lower=True, trans=1, unit_diagonal=True),
                               rtol=1e-12, atol=1e-12)
    # AutoLinearSolver picks Triangular for triangular tags
    x2 = lx.linear_solve(op, _t(b, device)).value.cpu().numpy()
    np.testing.assert_allclose(x2, x, rtol=1e-14, atol=1e-14)


def test_trsv_float32(device, rng):
    n = 300
    a = np.triu(rng.standard_normal((n, n)) / np.sqrt(n)) + 2 * np.eye(n)
    b = rng.standard_normal(n)
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    x = K.tri_solve(_t(a32, device, torch.float32), _t(b32, device, torch.float32), lower=False, trans=False,
                    unit_diag=False).cpu().numpy().astype(np.float64)
    ref = sla.solve_triangular(a32.astype(np.float64), b32.astype(np.float64))
    assert np.max(np.abs(x - ref)) <= 1e-4 * np.max(np.abs(ref))
