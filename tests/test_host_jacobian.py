"""JacobianLinearOperator (lineax/_operator.py:539-660) on the host.

The operator of fn(x) = a + B x + C x^2 at x has the closed-form Jacobian
J = B + C diag(2x) (the reference's tests/helpers.py:215-274 builds its
Jacobian operators the same way).  mv, batched mv, transpose, linearise and
materialise must reproduce J for jac = None / "fwd" / "bwd"; "bwd" must work
for a function that defines only a backward rule (no forward-mode
derivative), as the reference's custom_vjp case; pytree inputs are
flattened consistently.
"""

import pytest
import torch

import lineax_amd as lx


def _problem(n=6, m=None, seed=0):
    g = torch.Generator().manual_seed(seed)
    m = m or n
    a = torch.randn(m, dtype=torch.float64, generator=g)
    b = torch.randn(m, n, dtype=torch.float64, generator=g)
    c = torch.randn(m, n, dtype=torch.float64, generator=g)
    x = torch.randn(n, dtype=torch.float64, generator=g)
    fn = lambda x, _: a + b @ x + c @ (x ** 2)
    return fn, x, b + c * (2 * x)[None, :]


class _BackwardOnly(torch.autograd.Function):
    """a + B x + C x^2 with only a backward rule (no jvp)."""

    generate_vmap_rule = True

    @staticmethod
    def forward(x, a, b, c):
        return a + b @ x + c @ (x ** 2)

    @staticmethod
    def setup_context(ctx, inputs, output):
        x, a, b, c = inputs
        ctx.save_for_backward(x, b, c)

    @staticmethod
    def backward(ctx, g):
        x, b, c = ctx.saved_tensors
        return b.T @ g + 2 * (c.T @ g) * x, None, None, None


@pytest.mark.parametrize("jac", [None, "fwd", "bwd"])
def test_jacobian_operator_matches_closed_form(jac):
    fn, x, J = _problem()
    op = lx.JacobianLinearOperator(fn, x, None, (), jac=jac)
    g = torch.Generator().manual_seed(5)
    v = torch.randn(6, dtype=torch.float64, generator=g)
    V = torch.randn(2, 3, 6, dtype=torch.float64, generator=g)
    torch.testing.assert_close(op.mv(v), J @ v, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(op.mv(V), V @ J.T, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(op.as_matrix(), J, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(op.T.mv(v), J.T @ v, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(lx.linearise(op).mv(V), V @ J.T, rtol=1e-12, atol=1e-12)
    mat = lx.materialise(op)
    assert isinstance(mat, lx.MatrixLinearOperator)
    torch.testing.assert_close(mat.as_matrix(), J, rtol=1e-12, atol=1e-12)


def test_jacobian_operator_nonsquare_and_structures():
    fn, x, J = _problem(n=4, m=7, seed=2)
    op = lx.JacobianLinearOperator(fn, x)
    assert op.in_size() == 4 and op.out_size() == 7
    torch.testing.assert_close(op.as_matrix(), J, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        lx.JacobianLinearOperator(fn, x, jac="sideways")


def test_jacobian_operator_backward_only_function():
    g = torch.Generator().manual_seed(3)
    a = torch.randn(5, dtype=torch.float64, generator=g)
    b = torch.randn(5, 5, dtype=torch.float64, generator=g)
    c = torch.randn(5, 5, dtype=torch.float64, generator=g)
    x = torch.randn(5, dtype=torch.float64, generator=g)
    fn = lambda x, _: _BackwardOnly.apply(x, a, b, c)
    J = b + c * (2 * x)[None, :]
    op = lx.JacobianLinearOperator(fn, x, jac="bwd")
    v = torch.randn(5, dtype=torch.float64, generator=g)
    torch.testing.assert_close(op.mv(v), J @ v, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(op.T.mv(v), J.T @ v, rtol=1e-12, atol=1e-12)
    torch.testing.assert_close(op.as_matrix(), J, rtol=1e-12, atol=1e-12)
    # forward mode is unavailable for this function: jac="fwd" must fail
    with pytest.raises(Exception):
        lx.JacobianLinearOperator(fn, x, jac="fwd").mv(v)


def test_jacobian_operator_pytree_input_and_tags():
    g = torch.Generator().manual_seed(4)
    w = torch.randn(3, 3, dtype=torch.float64, generator=g)
    x = (torch.randn(3, dtype=torch.float64, generator=g), torch.randn(2, dtype=torch.float64, generator=g))
    fn = lambda x, _: (w @ x[0] + x[1].sum(), x[0][:2] * x[1])
    op = lx.JacobianLinearOperator(fn, x, tags=lx.symmetric_tag)
    assert lx.is_symmetric(op) and not lx.is_positive_semidefinite(op)
    m = op.as_matrix()
    assert m.shape == (5, 5)
    v = (torch.randn(3, dtype=torch.float64, generator=g), torch.randn(2, dtype=torch.float64, generator=g))
    out = op.mv(v)
    flat = torch.cat([out[0], out[1]])
    torch.testing.assert_close(flat, m @ torch.cat(v), rtol=1e-12, atol=1e-12)
    # the Hessian of a convex function is a PSD Jacobian (of its gradient)
    h = lx.JacobianLinearOperator(lambda x, _: 2 * x + torch.sin(x) * 0.1, torch.zeros(4, dtype=torch.float64),
                                  tags=lx.positive_semidefinite_tag)
    assert lx.is_positive_semidefinite(h) and lx.is_symmetric(h)
