"""Pins the dense oracle (oracle/lxo_dense.c) before it is trusted.

Each routine is checked against LAPACK through scipy (the library JAX calls
on CPU), against closed forms (PA = LU, U^T U = A, Q R = A, orthogonality),
and against the known-answer vectors of the reference tests.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lineax_oracle as O
from tests.helpers import well_conditioned


@pytest.mark.parametrize("n", [1, 2, 5, 32, 64, 100])
def test_getrf_matches_lapack(rng, n):
    a = rng.standard_normal((n, n))
    lu, piv, info = O.lu_factor(a)
    lu_ref, piv_ref = sla.lu_factor(a)
    assert info == 0
    np.testing.assert_array_equal(piv, piv_ref)
    np.testing.assert_allclose(lu, lu_ref, rtol=1e-10, atol=1e-12)


def test_getrf_reconstructs(rng):
    n = 40
    a = rng.standard_normal((n, n))
    lu, piv, _ = O.lu_factor(a)
    l = np.tril(lu, -1) + np.eye(n)
    u = np.triu(lu)
    pa = a.copy()
    for k in range(n):
        pa[[k, piv[k]]] = pa[[piv[k], k]]
    assert np.abs(pa - l @ u).max() <= 50 * n * np.finfo(float).eps * np.abs(a).max()


def test_getrf_first_max_tie_rule():
    # Two rows share the largest |a(i,0)|: idamax takes the first.
    a = np.array([[1.0, 2.0, 0.0], [-3.0, 1.0, 1.0], [3.0, 0.0, 2.0]])
    _, piv, _ = O.lu_factor(a)
    assert piv[0] == 1
    np.testing.assert_array_equal(piv, sla.lu_factor(a)[1])


@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("n", [1, 3, 33])
def test_getrs(rng, n, trans):
    a = well_conditioned(rng, n)
    b = rng.standard_normal(n)
    lu, piv, _ = O.lu_factor(a)
    x = O.lu_solve(lu, piv, b, trans)
    ref = np.linalg.solve(a.T if trans else a, b)
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-10)


def test_lu_zero_column_is_singular(rng):
    a = rng.standard_normal((4, 4))
    a[:, 2] = 0.0
    lu, piv, info = O.lu_factor(a)
    assert info == 3
    x = O.lu_solve(lu, piv, rng.standard_normal(4))
    assert O.linear_solve_results(x, np.ones(4), O.SUCCESSFUL) == O.SINGULAR


def test_lu_nonfinite_input_result(rng):
    a = well_conditioned(rng, 3)
    b = np.array([1.0, np.inf, 0.0])
    x, result = O.solve_dense("lu", a, b)
    assert result == O.NONFINITE_INPUT


def test_reference_kat_2x2():
    # lineax tests/test_solve.py:41-48 and 64-76 (the mixed-dtype variant
    # has the same values): [[1, 5], [-2, -2]] x = [3, 4] -> [-3.25, 1.25].
    a = np.array([[1.0, 5.0], [-2.0, -2.0]])
    b = np.array([3.0, 4.0])
    for kind in ("lu", "qr"):
        x, result = O.solve_dense(kind, a, b)
        assert result == O.SUCCESSFUL
        np.testing.assert_allclose(x, [-3.25, 1.25], rtol=1e-14)


@pytest.mark.parametrize("n", [1, 4, 32, 64, 65, 200])
def test_potrf_matches_lapack(rng, n):
    a = well_conditioned(rng, n, kind="psd")
    u, info = O.cho_factor(a)
    ref, lower = sla.cho_factor(a, lower=False)
    assert info == 0 and not lower
    np.testing.assert_allclose(u, np.triu(ref), rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(u.T @ u, a, rtol=1e-10, atol=1e-10 * np.abs(a).max())


def test_potrf_reads_upper_triangle_only(rng):
    a = well_conditioned(rng, 6, kind="psd")
    junk = a.copy()
    junk[np.tril_indices(6, -1)] = 1e300
    np.testing.assert_array_equal(O.cho_factor(junk)[0], O.cho_factor(a)[0])


def test_potrf_failure_is_nan(rng):
    a = well_conditioned(rng, 5, kind="psd")
    a[3, 3] = -10.0
    u, info = O.cho_factor(a)
    assert info > 0 and np.all(np.isnan(u))
    x = O.cho_solve(u, np.ones(5))
    assert O.linear_solve_results(x, np.ones(5), O.SUCCESSFUL) == O.SINGULAR


def test_cholesky_nsd_invariant(rng):
    # lineax/_solver/cholesky.py:58-59,75-76: NSD A solves as -solve(-A, b).
    a = well_conditioned(rng, 7, kind="nsd")
    b = rng.standard_normal(7)
    u, _ = O.cho_factor(a, negate=True)
    x = O.cho_solve(u, b, negate=True)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("shape", [(5, 5), (8, 3), (40, 17), (64, 64)])
def test_geqrf_matches_lapack(rng, shape):
    a = rng.standard_normal(shape)
    h, tau, transposed = O.qr_factor(a)
    assert not transposed
    (h_ref, tau_ref), _r = sla.qr(a, mode="raw")
    np.testing.assert_allclose(h, h_ref, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(tau, tau_ref, rtol=1e-10, atol=1e-12)


def test_ormqr_is_orthogonal(rng):
    a = rng.standard_normal((12, 5))
    h, tau, _ = O.qr_factor(a)
    c = rng.standard_normal(12)
    qtc = c.copy()
    O.lib().lxo_ormqr_vec(12, 5, O._dp(h), 5, O._dp(tau), O._dp(qtc), 1)
    back = qtc.copy()
    O.lib().lxo_ormqr_vec(12, 5, O._dp(h), 5, O._dp(tau), O._dp(back), 0)
    np.testing.assert_allclose(back, c, atol=1e-13)
    assert abs(np.linalg.norm(qtc) - np.linalg.norm(c)) < 1e-12
    # Q^T A = R
    q_full, _ = np.linalg.qr(a, mode="complete")
    r = np.triu(h)[:5]
    np.testing.assert_allclose(np.abs(np.linalg.qr(a)[1]), np.abs(r), atol=1e-12)
    del q_full


@pytest.mark.parametrize("shape", [(6, 3), (3, 6), (30, 12), (12, 30), (9, 9)])
def test_qr_solve_least_squares_and_min_norm(rng, shape):
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    x, result = O.solve_dense("qr", a, b)
    ref = np.linalg.lstsq(a, b, rcond=None)[0]
    assert result == O.SUCCESSFUL
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-10)


def test_reference_nonsquare_kats():
    # lineax tests/test_singular.py:113-145 (checked there against lstsq).
    a = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    y = np.array([3.0, 4.0])
    x, _ = O.solve_dense("qr", a, y)
    np.testing.assert_allclose(x, np.linalg.lstsq(a, y, rcond=None)[0], rtol=1e-12)
    a2 = np.array([[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]])
    y2 = np.array([3.0, 4.0, 5.0])
    x2, _ = O.solve_dense("qr", a2, y2)
    np.testing.assert_allclose(x2, np.linalg.lstsq(a2, y2, rcond=None)[0], rtol=1e-12)


def test_trsv_both_orientations(rng):
    r = np.triu(rng.standard_normal((6, 6))) + 4 * np.eye(6)
    b = rng.standard_normal(6)
    for trans, mat in ((0, r), (1, r.T)):
        x = b.copy()
        O.lib().lxo_trsv_upper(6, O._dp(np.ascontiguousarray(r)), 6, O._dp(x), trans)
        np.testing.assert_allclose(mat @ x, b, atol=1e-12)


def test_chol_uneven_driver(rng):
    mats, rhss, ns = [], [], [1, 7, 64, 130]
    for n in ns:
        mats.append(well_conditioned(rng, n, kind="psd").astype(np.float32))
        rhss.append(rng.standard_normal(n).astype(np.float32))
    from lineax_amd.workloads import pack_uneven

    host = pack_uneven(ns, mats, rhss)
    x, info = O.chol_uneven(host.n, host.a_offset, host.a, host.b_offset, host.b, threads=2)
    assert np.all(info == 0)
    for i, n in enumerate(ns):
        xi = x[host.b_offset[i]: host.b_offset[i] + n]
        ref = np.linalg.solve(mats[i].astype(np.float64), rhss[i].astype(np.float64))
        np.testing.assert_allclose(xi, ref, rtol=1e-9, atol=1e-9)


def test_backward_error_metric():
    a = np.array([[2.0, 0.0], [0.0, 4.0]])
    b = np.array([2.0, 4.0])
    assert O.normwise_backward_error(a, np.array([1.0, 1.0]), b) == 0.0
    eta = O.normwise_backward_error(a, np.array([1.0, 1.5]), b)
    # r = [0, -2]; |A|inf = 4, |x|inf = 1.5, |b|inf = 4 -> 2 / 10
    assert abs(eta - 0.2) < 1e-15


def test_lstsq_backward_error_metric():
    # A = [[1, 0], [0, 1], [1, 1]], b = [1, 2, 0]: A^T A = [[2, 1], [1, 2]],
    # A^T b = [1, 2] -> x* = [0, 1]; residual r = [1, 1, -1], A^T r = [0, 0]
    a = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    b = np.array([1.0, 2.0, 0.0])
    assert O.lstsq_backward_error(a, np.array([0.0, 1.0]), b) == 0.0
    # x = [1, 1]: r = [0, 1, -2], A^T r = [-2, -1]; |A|_1 = 2, |A|_inf = 2,
    # |x|_inf = 1, |b|_inf = 2 -> 2 / (2 * (2 + 2)) = 0.25
    assert abs(O.lstsq_backward_error(a, np.array([1.0, 1.0]), b) - 0.25) < 1e-15
    # the float64 lstsq minimiser of a random inconsistent problem sits at
    # rounding level, a float32-rounded one well above it
    rng = np.random.default_rng(4)
    a = rng.standard_normal((200, 60))
    b = rng.standard_normal(200)
    x = np.linalg.lstsq(a, b, rcond=None)[0]
    assert O.lstsq_backward_error(a, x, b) < 1e-14
    assert O.lstsq_backward_error(a, x.astype(np.float32), b) > 1e-10
    # a transposed-operand slip (A r in place of A^T r) cannot pass: shapes differ
    # unless square; for square A the two differ on a non-symmetric example
    sq = np.array([[1.0, 2.0], [0.0, 1.0]])
    bb = np.array([1.0, 1.0])
    xx = np.array([0.0, 0.0])  # r = b; A^T r = [1, 3], A r = [3, 1]
    assert abs(O.lstsq_backward_error(sq, xx, bb) - 3.0 / (3.0 * 1.0)) < 1e-15


def test_normal_solve_pins():
    rng = np.random.default_rng(9)
    # tall full rank: the least-squares solution
    a = rng.standard_normal((30, 8))
    b = rng.standard_normal(30)
    np.testing.assert_allclose(O.normal_solve(a, b), np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-10, atol=1e-12)
    # wide full rank: the minimum-norm solution, which solves A x = b exactly
    a = rng.standard_normal((6, 15))
    b = rng.standard_normal(6)
    x = O.normal_solve(a, b)
    np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(a @ x, b, rtol=1e-10, atol=1e-12)
    # the reference's non-square KATs (tests/test_singular.py:113-145)
    a1 = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    y1 = np.array([3.0, 4.0])
    a2 = np.array([[1.0, -2.0], [5.0, -2.0], [-1.0, 3.0]])
    y2 = np.array([3.0, 4.0, 5.0])
    for aa, yy in ((a1, y1), (a2, y2)):
        np.testing.assert_allclose(O.normal_solve(aa, yy), np.linalg.lstsq(aa, yy, rcond=None)[0], rtol=1e-12,
                                   atol=1e-12)
    # square: the solution of A x = b (a transposed Gram matrix would fail this)
    a = np.array([[2.0, 1.0], [0.0, 3.0]])
    b = np.array([5.0, 6.0])
    np.testing.assert_allclose(O.normal_solve(a, b), [1.5, 2.0], rtol=1e-13)


def test_svd_solve_pins():
    rng = np.random.default_rng(11)
    # full rank, tall and wide: lstsq / minimum norm
    for shape in ((20, 7), (7, 20), (9, 9)):
        a = rng.standard_normal(shape)
        b = rng.standard_normal(shape[0])
        x, rank = O.svd_solve(a, b)
        assert rank == min(shape)
        np.testing.assert_allclose(x, np.linalg.lstsq(a, b, rcond=None)[0], rtol=1e-10, atol=1e-12)
    # rank deficient: the pseudoinverse solution (numpy.linalg.pinv)
    a = rng.standard_normal((12, 3)) @ rng.standard_normal((3, 8))
    b = rng.standard_normal(12)
    x, rank = O.svd_solve(a, b)
    assert rank == 3
    np.testing.assert_allclose(x, np.linalg.pinv(a) @ b, rtol=1e-8, atol=1e-10)
    # the zero matrix has rank 0 and solution 0 (the strict > cutoff)
    x, rank = O.svd_solve(np.zeros((3, 4)), np.ones(3))
    assert rank == 0 and np.all(x == 0)
    # the reference's non-square KATs (tests/test_singular.py:100-145)
    a1 = np.array([[1.0, 5.0, -1.0], [-2.0, -2.0, 3.0]])
    np.testing.assert_allclose(O.svd_solve(a1, [3.0, 4.0])[0], np.linalg.lstsq(a1, [3.0, 4.0], rcond=None)[0],
                               rtol=1e-12, atol=1e-12)
