"""Derivatives through linear_solve on the GPU against closed forms.

lineax's JVP rule (lineax/_solve.py:151-268) gives, for the solution
x = A^+ b of a full-rank operator,

  x' = A^+ (b' - A' x)                                  (square)
     + A^+ A^+T A'^T (b - A x)                          (tall: A^+ A = I)
     + (I - A^+ A) A'^T A^+T x                          (wide: A A^+ = I)

and its transpose (lineax/_solve.py:286-317) the reverse mode.  The
expected tangents are the textbook derivatives of the least-squares and
minimum-norm solutions, computed with numpy in float64
(tests.helpers.dx_closed_form, itself pinned against central differences of
numpy's solutions in test_host_derivative_forms.py), so a dropped term or a
transposed operand fails.  Forward mode is checked for every solver of this
path that lineax's tests/test_jvp.py:46-120 covers (LU, Cholesky, QR, CG,
GMRES), and the non-square branches for QR both ways (lineax's
tests/test_singular.py:162-255).
"""

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

import lineax_amd as lx
from tests.helpers import dx_closed_form as _dx_closed_form
from tests.helpers import lstsq as _lstsq
from tests.helpers import well_conditioned

pytestmark = pytest.mark.gpu


def _t(x, device):
    return torch.as_tensor(np.asarray(x), dtype=torch.float64, device=device)


def _jvp(device, a, b, da, db, solver, tags=()):
    with fwAD.dual_level():
        A = fwAD.make_dual(_t(a, device), _t(da, device))
        B = fwAD.make_dual(_t(b, device), _t(db, device))
        x = lx.linear_solve(lx.MatrixLinearOperator(A, tags), B, solver).value
        primal, tangent = fwAD.unpack_dual(x)
    return primal.cpu().numpy(), tangent.cpu().numpy()


@pytest.mark.parametrize("name", ["lu", "cholesky", "qr", "cg", "gmres"])
def test_jvp_square_every_solver(device, rng, name):
    n = 12
    psd = name in ("cholesky", "cg")
    a = well_conditioned(rng, n, kind="psd" if psd else "general")
    b = rng.standard_normal(n)
    da = rng.standard_normal((n, n))
    if psd:
        da = da + da.T  # a tangent that keeps the operator symmetric
    db = rng.standard_normal(n)
    solver = {"lu": lx.LU(), "cholesky": lx.Cholesky(), "qr": lx.QR(),
              "cg": lx.CG(rtol=1e-13, atol=1e-13), "gmres": lx.GMRES(rtol=1e-13, atol=1e-13, restart=n)}[name]
    tags = lx.positive_semidefinite_tag if psd else ()
    x, dx = _jvp(device, a, b, da, db, solver, tags)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(dx, _dx_closed_form(a, b, da, db), rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("shape", [(14, 6), (40, 25), (5, 11), (20, 33)])
def test_jvp_qr_nonsquare(device, shape):
    rng = np.random.default_rng(shape[0] * 100 + shape[1])
    a = rng.standard_normal(shape)
    b, da, db = rng.standard_normal(shape[0]), rng.standard_normal(shape), rng.standard_normal(shape[0])
    x, dx = _jvp(device, a, b, da, db, lx.QR())
    np.testing.assert_allclose(x, _lstsq(a, b), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(dx, _dx_closed_form(a, b, da, db), rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("shape", [(14, 6), (5, 11)])
def test_jvp_qr_nonsquare_tangent_of_b_only_and_a_only(device, shape):
    # each branch separately: b' alone (only A^+ b' survives) and A' alone
    rng = np.random.default_rng(7 + shape[0])
    a = rng.standard_normal(shape)
    b = rng.standard_normal(shape[0])
    zeros_a, zeros_b = np.zeros(shape), np.zeros(shape[0])
    da, db = rng.standard_normal(shape), rng.standard_normal(shape[0])
    _, dx_b = _jvp(device, a, b, zeros_a, db, lx.QR())
    np.testing.assert_allclose(dx_b, np.linalg.pinv(a) @ db, rtol=1e-8, atol=1e-9)
    _, dx_a = _jvp(device, a, b, da, zeros_b, lx.QR())
    np.testing.assert_allclose(dx_a, _dx_closed_form(a, b, da, zeros_b), rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("shape", [(14, 6), (40, 25), (5, 11), (20, 33), (9, 9)])
def test_vjp_qr_matches_adjoint_of_closed_form(device, shape):
    # reverse mode: <grad_A, dA> + <grad_b, db> = g . x'(dA, db) for any direction
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    a = rng.standard_normal(shape) + (3 * np.eye(*shape) if shape[0] == shape[1] else 0)
    b, g = rng.standard_normal(shape[0]), rng.standard_normal(shape[1])
    A = _t(a, device).requires_grad_(True)
    B = _t(b, device).requires_grad_(True)
    x = lx.linear_solve(lx.MatrixLinearOperator(A), B, lx.QR()).value
    (x * _t(g, device)).sum().backward()
    ga, gb = A.grad.cpu().numpy(), B.grad.cpu().numpy()
    for _ in range(3):
        da, db = rng.standard_normal(shape), rng.standard_normal(shape[0])
        lhs = np.sum(ga * da) + gb @ db
        rhs = g @ _dx_closed_form(a, b, da, db)
        assert abs(lhs - rhs) <= 1e-8 * max(1.0, abs(rhs)), (lhs, rhs)
    # the b-gradient alone is A^+T g
    np.testing.assert_allclose(gb, np.linalg.pinv(a).T @ g, rtol=1e-8, atol=1e-9)


def test_jvp_batched_leading_dims(device, rng):
    # vmap analogue: a [3, n, n] operator and [3, n] vectors, tangent per system
    n = 8
    a = np.stack([well_conditioned(rng, n) for _ in range(3)])
    b, da, db = rng.standard_normal((3, n)), rng.standard_normal((3, n, n)), rng.standard_normal((3, n))
    x, dx = _jvp(device, a, b, da, db, lx.LU())
    for i in range(3):
        np.testing.assert_allclose(dx[i], _dx_closed_form(a[i], b[i], da[i], db[i]), rtol=1e-8, atol=1e-9)
