"""Linear operators and their structural predicates (lineax/_operator.py).

Operators on the solver hot path: ``MatrixLinearOperator`` (whose ``mv`` is
the GEMV kernel), ``FunctionLinearOperator``, ``TaggedLinearOperator`` and
``IdentityLinearOperator``, plus the scalar/negation/sum/composition wrappers
lineax builds from operator arithmetic.  The grid stencil operators used by
the Krylov configurations live in ``lineax_amd.stencil``.

Batching: a ``MatrixLinearOperator`` may hold a stack of matrices
``[*batch, m, n]``; its structures describe one member of the stack.  Vectors
given to ``mv`` (and to ``linear_solve``) may carry leading batch dimensions
that broadcast against the operator's.  ``FunctionLinearOperator`` functions
must accept such leading dimensions.
"""

from __future__ import annotations

import abc
import functools as ft
import warnings
from typing import Callable

import torch

from . import _kernels as K
from ._structure import (
    TensorStruct,
    batch_shape_of,
    default_floating_dtype,
    ravel,
    structure_equal,
    structure_of,
    tree_flatten,
    tree_leaves,
    tree_map,
    tree_unflatten,
    unravel,
)
from ._tags import (
    as_tag_set,
    diagonal_tag,
    lower_triangular_tag,
    negative_semidefinite_tag,
    positive_semidefinite_tag,
    symmetric_tag,
    transpose_tags,
    tridiagonal_tag,
    unit_diagonal_tag,
    upper_triangular_tag,
)

_sentinel = object()


def _size(struct) -> int:
    return sum(s.size for s in tree_leaves(struct))


class AbstractLinearOperator(abc.ABC):
    """Base class of linear operators X -> Y between pytrees of tensors.

    Supports ``op1 + op2``, ``op1 - op2``, ``op1 @ op2``, ``op * s``,
    ``op / s`` and ``-op`` (lineax/_operator.py:70-229).
    """

    def _check_square_structures(self):
        # lineax/_operator.py:96-110: symmetric/definite operators need equal
        # input and output structures.
        if is_symmetric(self) or is_positive_semidefinite(self) or is_negative_semidefinite(self):
            if not structure_equal(self.in_structure(), self.out_structure()):
                raise ValueError(
                    "Symmetric/Hermitian matrices must have matching input and output structures. "
                    f"Got input structure {self.in_structure()} and output structure {self.out_structure()}."
                )

    @abc.abstractmethod
    def mv(self, vector):
        """Matrix-vector product with a vector of structure ``in_structure()``."""

    @abc.abstractmethod
    def as_matrix(self) -> torch.Tensor:
        """Materialises the operator as a ``[*batch, out_size, in_size]`` tensor."""

    @abc.abstractmethod
    def transpose(self) -> "AbstractLinearOperator":
        """The transposed operator."""

    @abc.abstractmethod
    def in_structure(self):
        """Pytree of TensorStruct describing one input vector."""

    @abc.abstractmethod
    def out_structure(self):
        """Pytree of TensorStruct describing one output vector."""

    @property
    def batch_shape(self) -> tuple:
        return ()

    @property
    def device(self):
        return None

    def in_size(self) -> int:
        return _size(self.in_structure())

    def out_size(self) -> int:
        return _size(self.out_structure())

    @property
    def T(self) -> "AbstractLinearOperator":
        return self.transpose()

    # Tensor leaves, for stop-gradient and autodiff plumbing.
    def _leaves(self) -> list:
        return []

    def _with_leaves(self, leaves) -> "AbstractLinearOperator":
        return self

    def __add__(self, other):
        if not isinstance(other, AbstractLinearOperator):
            raise ValueError("Can only add AbstractLinearOperators together.")
        return AddLinearOperator(self, other)

    def __sub__(self, other):
        if not isinstance(other, AbstractLinearOperator):
            raise ValueError("Can only add AbstractLinearOperators together.")
        return AddLinearOperator(self, -other)

    def __mul__(self, other):
        if isinstance(other, torch.Tensor) and other.dim() != 0:
            raise ValueError("Can only multiply AbstractLinearOperators by scalars.")
        return MulLinearOperator(self, other)

    def __rmul__(self, other):
        return self * other

    def __matmul__(self, other):
        if not isinstance(other, AbstractLinearOperator):
            raise ValueError("Can only compose AbstractLinearOperators together.")
        return ComposedLinearOperator(self, other)

    def __truediv__(self, other):
        if isinstance(other, torch.Tensor) and other.dim() != 0:
            raise ValueError("Can only divide AbstractLinearOperators by scalars.")
        return DivLinearOperator(self, other)

    def __neg__(self):
        return NegLinearOperator(self)


# ---------------------------------------------------------------------------
# MatrixLinearOperator
# ---------------------------------------------------------------------------


class _MatVec(torch.autograd.Function):
    """y = A x on the GEMV kernel, differentiable in A and x."""

    @staticmethod
    def forward(a, x):
        return K.gemv(a, x, trans=False)

    @staticmethod
    def setup_context(ctx, inputs, output):
        a, x = inputs
        ctx.save_for_backward(a, x)
        ctx.save_for_forward(a, x)

    @staticmethod
    def backward(ctx, g):
        a, x = ctx.saved_tensors
        ga = gx = None
        if ctx.needs_input_grad[0]:
            ga = g.unsqueeze(-1) * x.unsqueeze(-2)
            ga = _sum_to(ga, a.shape)
        if ctx.needs_input_grad[1]:
            gx = _sum_to(K.gemv(a, g.contiguous(), trans=True), x.shape)
        return ga, gx

    @staticmethod
    def jvp(ctx, da, dx):
        a, x = ctx.saved_tensors
        out = None
        if da is not None:
            out = K.gemv(da.contiguous(), x, trans=False)
        if dx is not None:
            t = K.gemv(a, dx.contiguous(), trans=False)
            out = t if out is None else out + t
        return out


def _sum_to(t: torch.Tensor, shape) -> torch.Tensor:
    """Sums broadcast leading dims of t down to ``shape``."""
    while t.dim() > len(shape):
        t = t.sum(0)
    for i, s in enumerate(shape):
        if s == 1 and t.shape[i] != 1:
            t = t.sum(i, keepdim=True)
    return t


def matvec(matrix: torch.Tensor, vector: torch.Tensor) -> torch.Tensor:
    """matrix @ vector on the GEMV kernel (autodiff-aware)."""
    if matrix.requires_grad or vector.requires_grad or _is_dual(matrix) or _is_dual(vector):
        return _MatVec.apply(matrix, vector)
    return K.gemv(matrix, vector, trans=False)


def _is_dual(t) -> bool:
    try:
        import torch.autograd.forward_ad as fwAD

        return isinstance(t, torch.Tensor) and fwAD.unpack_dual(t).tangent is not None
    except Exception:  # pragma: no cover - forward AD unavailable
        return False


class MatrixLinearOperator(AbstractLinearOperator):
    """Wraps a 2-D tensor (or a stack ``[*batch, m, n]``) as a linear operator.

    ``mv`` takes vectors of shape ``(n,)`` (plus batch dims) to ``(m,)``
    (lineax/_operator.py:232-285).
    """

    def __init__(self, matrix: torch.Tensor, tags=()):
        if not isinstance(matrix, torch.Tensor):
            matrix = torch.as_tensor(matrix)
        if matrix.dim() < 2:
            raise ValueError("`MatrixLinearOperator(matrix=...)` should be 2-dimensional.")
        if not (matrix.is_floating_point() or matrix.is_complex()):
            matrix = matrix.to(torch.float32)
        self.matrix = matrix
        self.tags = as_tag_set(tags)
        self._check_square_structures()

    def mv(self, vector):
        sparse = _try_sparse_materialise(self)
        if sparse is not self:
            return sparse.mv(vector)
        return matvec(self.matrix, vector)

    def as_matrix(self):
        return self.matrix

    def transpose(self):
        if is_symmetric(self):
            return self
        return MatrixLinearOperator(self.matrix.mT, transpose_tags(self.tags))

    def in_structure(self):
        return TensorStruct((self.matrix.shape[-1],), self.matrix.dtype)

    def out_structure(self):
        return TensorStruct((self.matrix.shape[-2],), self.matrix.dtype)

    @property
    def batch_shape(self):
        return tuple(self.matrix.shape[:-2])

    @property
    def device(self):
        return self.matrix.device

    def _leaves(self):
        return [self.matrix]

    def _with_leaves(self, leaves):
        out = object.__new__(MatrixLinearOperator)
        out.matrix = leaves[0]
        out.tags = self.tags
        return out

    def __repr__(self):
        return f"MatrixLinearOperator(shape={tuple(self.matrix.shape)}, dtype={self.matrix.dtype}, tags={set(self.tags)})"


class _DiagonalOperator(AbstractLinearOperator):
    """Elementwise scaling; the sparse form of a diagonal-tagged matrix."""

    def __init__(self, diag: torch.Tensor, struct):
        self.diag = diag
        self.struct = struct

    def mv(self, vector):
        return self.diag * vector

    def as_matrix(self):
        return torch.diag_embed(self.diag)

    def transpose(self):
        return self

    def in_structure(self):
        return self.struct

    def out_structure(self):
        return self.struct


class _TridiagonalOperator(AbstractLinearOperator):
    """Tridiagonal product; the sparse form of a tridiagonal-tagged matrix."""

    def __init__(self, diag, lower, upper, struct):
        self.diag, self.lower, self.upper, self.struct = diag, lower, upper, struct

    def mv(self, vector):
        out = self.diag * vector
        out = out.clone()
        out[..., :-1] += self.upper * vector[..., 1:]
        out[..., 1:] += self.lower * vector[..., :-1]
        return out

    def as_matrix(self):
        return (
            torch.diag_embed(self.diag)
            + torch.diag_embed(self.upper, offset=1)
            + torch.diag_embed(self.lower, offset=-1)
        )

    def transpose(self):
        return _TridiagonalOperator(self.diag, self.upper, self.lower, self.struct)

    def in_structure(self):
        return self.struct

    def out_structure(self):
        return self.struct


def _try_sparse_materialise(operator):
    """Diagonal/tridiagonal-tagged matrices multiply in their sparse form.

    lineax/_operator.py:1318-1337.
    """
    if isinstance(operator, MatrixLinearOperator):
        tags = operator.tags
        if diagonal_tag in tags:
            return _DiagonalOperator(torch.diagonal(operator.matrix, dim1=-2, dim2=-1), operator.in_structure())
        if tridiagonal_tag in tags:
            m = operator.matrix
            return _TridiagonalOperator(
                torch.diagonal(m, dim1=-2, dim2=-1),
                torch.diagonal(m, offset=-1, dim1=-2, dim2=-1),
                torch.diagonal(m, offset=1, dim1=-2, dim2=-1),
                operator.in_structure(),
            )
    return operator


# ---------------------------------------------------------------------------
# FunctionLinearOperator
# ---------------------------------------------------------------------------


def _as_structure(x):
    if isinstance(x, TensorStruct):
        return x
    if isinstance(x, torch.Tensor):
        return TensorStruct(tuple(x.shape), x.dtype)
    return tree_map(lambda leaf: leaf if isinstance(leaf, TensorStruct) else TensorStruct(tuple(leaf.shape), leaf.dtype), x)


def _inexact_structure(struct):
    def fix(s):
        if s.dtype.is_floating_point or s.dtype.is_complex:
            return s
        return TensorStruct(s.shape, default_floating_dtype())

    return tree_map(fix, struct)


class FunctionLinearOperator(AbstractLinearOperator):
    """Wraps a linear function ``fn: X -> Y`` (lineax/_operator.py:664-732).

    ``fn`` must accept leading batch dimensions on its input.  The output
    structure is inferred by evaluating ``fn`` on meta tensors, or on zeros
    when that fails; pass ``output_structure`` to skip the inference.
    """

    def __init__(self, fn: Callable, input_structure, tags=(), output_structure=None, device=None):
        self.fn = fn
        self.input_structure = _inexact_structure(_as_structure(input_structure))
        self.tags = as_tag_set(tags)
        self._out = None if output_structure is None else _as_structure(output_structure)
        self._device = device
        self._check_square_structures()

    def mv(self, vector):
        return self.fn(vector)

    def as_matrix(self):
        return _materialise_by_columns(self)

    def transpose(self):
        if is_symmetric(self):
            return self
        return FunctionLinearOperator(
            _LinearTranspose(self.fn, self.in_structure()),
            self.out_structure(),
            transpose_tags(self.tags),
            output_structure=self.in_structure(),
            device=self._device,
        )

    def in_structure(self):
        return self.input_structure

    def out_structure(self):
        if self._out is None:
            self._out = _infer_out_structure(self.fn, self.input_structure, self._device)
        return self._out

    @property
    def device(self):
        return self._device


class _PerVector:
    """Applies a function of one unbatched vector (pytree) to a vector with
    leading batch dimensions, through torch.func.vmap over the flattened batch."""

    def __init__(self, fn, in_struct):
        self.fn = fn
        self.in_struct = in_struct

    def __call__(self, vector):
        batch = batch_shape_of(vector, self.in_struct)
        if not batch:
            return self.fn(vector)
        nb = 1
        for d in batch:
            nb *= d
        flat = tree_map(lambda t, s: t.reshape(nb, *s.shape), vector, self.in_struct)
        out = torch.func.vmap(self.fn)(flat)
        return tree_map(lambda t: t.reshape(*batch, *t.shape[1:]), out)


class JacobianLinearOperator(AbstractLinearOperator):
    """The Jacobian (d fn / d x)(x, args) of ``fn: (x, args) -> y`` at ``x``,
    never materialised unless asked (lineax/_operator.py:539-660).

    ``mv`` is a Jacobian-vector product: ``torch.func.jvp`` for ``jac=None``
    or ``"fwd"``; for ``"bwd"`` the transpose of ``torch.func.vjp`` (a
    vector-Jacobian product of the linear map u -> J^T u), which needs only
    reverse-mode derivatives of ``fn`` (lineax/_operator.py:621-641), or the
    vjp itself when the operator is tagged symmetric.  ``transpose`` is the
    vjp as a FunctionLinearOperator (lineax/_operator.py:646-654).
    ``linearise`` caches the primal pass (``torch.func.linearize``) and
    ``materialise`` forms the Jacobian with ``jacfwd`` / ``jacrev`` by
    lineax's heuristic (lineax/_misc.py:41-55).  ``fn`` must be built from
    differentiable torch operations.  Derivatives of a solve with respect to
    the point ``x`` or ``args`` are not propagated (they are constants of the
    operator); derivatives with respect to the right-hand side are.
    """

    def __init__(self, fn: Callable, x, args=None, tags=(), jac=None):
        if jac not in (None, "fwd", "bwd"):
            raise ValueError(
                "`jac` argument of `JacobianLinearOperator` should be either `'fwd'`, `'bwd'`, or `None`."
            )
        self.fn = fn
        self.x = tree_map(_inexact_leaf, x)
        self.args = args
        self.tags = as_tag_set(tags)
        self.jac = jac
        self._in = structure_of(self.x)
        self._out = None
        self._check_square_structures()

    def _f(self, x):
        return self.fn(x, self.args)

    def _jvp(self, v):
        return torch.func.jvp(self._f, (self.x,), (v,))[1]

    def _jv_from_vjp(self, v):
        y, vjp_fn = torch.func.vjp(self._f, self.x)
        if is_symmetric(self):
            return vjp_fn(v)[0]
        zeros = tree_map(torch.zeros_like, y)
        return torch.func.vjp(lambda u: vjp_fn(u)[0], zeros)[1](v)[0]

    def mv(self, vector):
        one = self._jv_from_vjp if self.jac == "bwd" else self._jvp
        return _PerVector(one, self.in_structure())(vector)

    def as_matrix(self):
        return _jacobian_matrix(self)

    def transpose(self):
        if is_symmetric(self):
            return self
        _, vjp_fn = torch.func.vjp(self._f, self.x)
        return FunctionLinearOperator(
            _PerVector(lambda y: vjp_fn(y)[0], self.out_structure()),
            self.out_structure(),
            transpose_tags(self.tags),
            output_structure=self.in_structure(),
            device=self.device,
        )

    def in_structure(self):
        return self._in

    def out_structure(self):
        if self._out is None:
            with torch.no_grad():
                self._out = structure_of(self._f(self.x))
        return self._out

    @property
    def device(self):
        leaves = tree_leaves(self.x)
        return leaves[0].device if leaves else None

    def __repr__(self):
        return f"JacobianLinearOperator(in={self._in}, jac={self.jac!r}, tags={set(self.tags)})"


def _inexact_leaf(t):
    t = torch.as_tensor(t)
    if not (t.is_floating_point() or t.is_complex()):
        t = t.to(default_floating_dtype())
    return t.detach()


def _jacobian_matrix(op: "JacobianLinearOperator") -> torch.Tensor:
    """[out_size, in_size] Jacobian of op.fn at op.x on flattened vectors."""
    in_struct, out_struct = op.in_structure(), op.out_structure()
    in_size, out_size = op.in_size(), op.out_size()

    def flat_fn(xf):
        return ravel(op._f(unravel(xf, in_struct, ())), out_struct, ())

    x_flat = ravel(op.x, in_struct, ())
    if op.jac is None:
        fwd = in_size < 100 or in_size <= 1.5 * out_size
    else:
        fwd = op.jac == "fwd"
    jac = (torch.func.jacfwd if fwd else torch.func.jacrev)(flat_fn)(x_flat)
    return jac.reshape(out_size, in_size)


class _LinearTranspose:
    """y -> (d fn / d x)^T y for linear fn: the vector-Jacobian product at zero.

    Stands in for jax.linear_transpose (lineax/_operator.py:711-723).
    """

    def __init__(self, fn, in_struct):
        self.fn = fn
        self.in_struct = in_struct

    def __call__(self, y):
        out_leaves = tree_leaves(y)
        dev = out_leaves[0].device
        ystruct_batch = None
        # batch dims of y relative to fn's output structure are those of the input too
        in_leaves, spec = tree_flatten(self.in_struct)
        probe = self.fn(tree_unflatten([torch.zeros(s.shape, dtype=s.dtype, device=dev) for s in in_leaves], spec))
        probe_leaves = tree_leaves(probe)
        k = probe_leaves[0].dim()
        ystruct_batch = tuple(out_leaves[0].shape[: out_leaves[0].dim() - k])
        with torch.enable_grad():
            xs = [
                torch.zeros((*ystruct_batch, *s.shape), dtype=s.dtype, device=dev, requires_grad=True)
                for s in in_leaves
            ]
            out = self.fn(tree_unflatten(xs, spec))
            grads = torch.autograd.grad(
                tree_leaves(out), xs, grad_outputs=out_leaves, allow_unused=True, create_graph=torch.is_grad_enabled()
            )
        grads = [torch.zeros_like(x) if g is None else g for g, x in zip(grads, xs)]
        return tree_unflatten(grads, spec)


def _infer_out_structure(fn, in_struct, device):
    leaves, spec = tree_flatten(in_struct)
    try:
        out = fn(tree_unflatten([torch.empty(s.shape, dtype=s.dtype, device="meta") for s in leaves], spec))
        return structure_of(out)
    except Exception:
        pass
    dev = device if device is not None else ("cuda" if torch.cuda.is_available() else "cpu")
    out = fn(tree_unflatten([torch.zeros(s.shape, dtype=s.dtype, device=dev) for s in leaves], spec))
    return structure_of(out)


def _materialise_by_columns(op: AbstractLinearOperator) -> torch.Tensor:
    """[out_size, in_size] matrix of ``op`` from its action on unit vectors."""
    in_struct = op.in_structure()
    n = op.in_size()
    leaves = tree_leaves(in_struct)
    dtype = leaves[0].dtype if leaves else default_floating_dtype()
    dev = op.device if op.device is not None else ("cuda" if torch.cuda.is_available() else "cpu")
    eye = torch.eye(n, dtype=dtype, device=dev)
    cols = op.mv(unravel(eye, in_struct, (n,)))
    flat = ravel(cols, op.out_structure(), (n,))
    return flat.mT


# ---------------------------------------------------------------------------
# Identity and wrappers
# ---------------------------------------------------------------------------


class IdentityLinearOperator(AbstractLinearOperator):
    """The identity X -> X, or the pad/truncate map between equal-size structures.

    lineax/_operator.py:736-823.
    """

    def __init__(self, input_structure, output_structure=_sentinel):
        if output_structure is _sentinel:
            output_structure = input_structure
        self.input_structure = _inexact_structure(_as_structure(input_structure))
        self.output_structure = _inexact_structure(_as_structure(output_structure))

    def mv(self, vector):
        batch = batch_shape_of(vector, self.input_structure)
        if structure_equal(self.input_structure, self.output_structure):
            return vector
        flat = ravel(vector, self.input_structure, batch)
        out = self.out_size()
        if flat.shape[-1] < out:
            pad = torch.zeros((*batch, out - flat.shape[-1]), dtype=flat.dtype, device=flat.device)
            flat = torch.cat([flat, pad], dim=-1)
        else:
            flat = flat[..., :out]
        return unravel(flat, self.output_structure, batch)

    def as_matrix(self):
        leaves = tree_leaves(self.input_structure)
        dtype = leaves[0].dtype if leaves else default_floating_dtype()
        return torch.eye(self.out_size(), self.in_size(), dtype=dtype)

    def transpose(self):
        return IdentityLinearOperator(self.output_structure, self.input_structure)

    def in_structure(self):
        return self.input_structure

    def out_structure(self):
        return self.output_structure

    @property
    def tags(self):
        return frozenset()


class TaggedLinearOperator(AbstractLinearOperator):
    """Another operator plus declared tags (lineax/_operator.py:890-942)."""

    def __init__(self, operator: AbstractLinearOperator, tags):
        self.operator = operator
        self.tags = as_tag_set(tags)
        self._check_square_structures()

    def mv(self, vector):
        return self.operator.mv(vector)

    def as_matrix(self):
        return self.operator.as_matrix()

    def transpose(self):
        return TaggedLinearOperator(self.operator.transpose(), transpose_tags(self.tags))

    def in_structure(self):
        return self.operator.in_structure()

    def out_structure(self):
        return self.operator.out_structure()

    @property
    def batch_shape(self):
        return self.operator.batch_shape

    @property
    def device(self):
        return self.operator.device

    def _leaves(self):
        return self.operator._leaves()

    def _with_leaves(self, leaves):
        return TaggedLinearOperator(self.operator._with_leaves(leaves), self.tags)


class _Unary(AbstractLinearOperator):
    def in_structure(self):
        return self.operator.in_structure()

    def out_structure(self):
        return self.operator.out_structure()

    @property
    def batch_shape(self):
        return self.operator.batch_shape

    @property
    def device(self):
        return self.operator.device

    def _leaves(self):
        return self.operator._leaves()


class NegLinearOperator(_Unary):
    """-operator (lineax/_operator.py:1064-1090)."""

    def __init__(self, operator):
        self.operator = operator

    def mv(self, vector):
        return tree_map(torch.neg, self.operator.mv(vector))

    def as_matrix(self):
        return -self.operator.as_matrix()

    def transpose(self):
        return NegLinearOperator(self.operator.transpose())

    def _with_leaves(self, leaves):
        return NegLinearOperator(self.operator._with_leaves(leaves))


class MulLinearOperator(_Unary):
    """operator * scalar (lineax/_operator.py:1031-1061)."""

    def __init__(self, operator, scalar):
        self.operator = operator
        self.scalar = scalar

    def mv(self, vector):
        return tree_map(lambda y: y * self.scalar, self.operator.mv(vector))

    def as_matrix(self):
        return self.operator.as_matrix() * self.scalar

    def transpose(self):
        return MulLinearOperator(self.operator.transpose(), self.scalar)

    def _with_leaves(self, leaves):
        return MulLinearOperator(self.operator._with_leaves(leaves), self.scalar)


class DivLinearOperator(_Unary):
    """operator / scalar (lineax/_operator.py:1093-1122)."""

    def __init__(self, operator, scalar):
        self.operator = operator
        self.scalar = scalar

    def mv(self, vector):
        return tree_map(lambda y: y / self.scalar, self.operator.mv(vector))

    def as_matrix(self):
        return self.operator.as_matrix() / self.scalar

    def transpose(self):
        return DivLinearOperator(self.operator.transpose(), self.scalar)

    def _with_leaves(self, leaves):
        return DivLinearOperator(self.operator._with_leaves(leaves), self.scalar)


class AddLinearOperator(AbstractLinearOperator):
    """operator1 + operator2 (lineax/_operator.py:989-1028)."""

    def __init__(self, operator1, operator2):
        if not structure_equal(operator1.in_structure(), operator2.in_structure()):
            raise ValueError("Incompatible linear operator structures")
        if not structure_equal(operator1.out_structure(), operator2.out_structure()):
            raise ValueError("Incompatible linear operator structures")
        self.operator1, self.operator2 = operator1, operator2

    def mv(self, vector):
        return tree_map(torch.add, self.operator1.mv(vector), self.operator2.mv(vector))

    def as_matrix(self):
        return self.operator1.as_matrix() + self.operator2.as_matrix()

    def transpose(self):
        return AddLinearOperator(self.operator1.transpose(), self.operator2.transpose())

    def in_structure(self):
        return self.operator1.in_structure()

    def out_structure(self):
        return self.operator1.out_structure()

    @property
    def device(self):
        return self.operator1.device or self.operator2.device


class ComposedLinearOperator(AbstractLinearOperator):
    """operator1 @ operator2 (lineax/_operator.py:1125-1192)."""

    def __init__(self, operator1, operator2):
        if not structure_equal(operator1.in_structure(), operator2.out_structure()):
            raise ValueError("Incompatible linear operator structures")
        self.operator1, self.operator2 = operator1, operator2

    def mv(self, vector):
        return self.operator1.mv(self.operator2.mv(vector))

    def as_matrix(self):
        m1, m2 = self.operator1.as_matrix(), self.operator2.as_matrix()
        if (m1.is_cuda and m1.dtype == torch.float32 and m2.dtype == torch.float32
                and not (m1.requires_grad or m2.requires_grad)):
            # one matrix-core launch over the whole batch (e.g. Normal's
            # Gram matrix A^T A)
            from . import _kernels as K

            return K.gemm(m1, m2)
        return m1 @ m2

    def transpose(self):
        return ComposedLinearOperator(self.operator2.transpose(), self.operator1.transpose())

    def in_structure(self):
        return self.operator2.in_structure()

    def out_structure(self):
        return self.operator1.out_structure()

    @property
    def device(self):
        return self.operator1.device or self.operator2.device


# ---------------------------------------------------------------------------
# Structural predicates (lineax/_operator.py:1587-2295)
# ---------------------------------------------------------------------------


def _has_real_dtype(op) -> bool:
    leaves = tree_leaves((op.in_structure(), op.out_structure()))
    return not any(s.dtype.is_complex for s in leaves)


def _tagged(op) -> frozenset:
    return getattr(op, "tags", frozenset())


@ft.singledispatch
def is_symmetric(operator) -> bool:
    """Whether ``operator`` is declared symmetric (symmetric/diagonal tags, or
    PSD/NSD with a real dtype)."""
    raise NotImplementedError(f"`is_symmetric` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_diagonal(operator) -> bool:
    raise NotImplementedError(f"`is_diagonal` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_tridiagonal(operator) -> bool:
    raise NotImplementedError(f"`is_tridiagonal` has not been implemented for {type(operator)}")


@ft.singledispatch
def has_unit_diagonal(operator) -> bool:
    raise NotImplementedError(f"`has_unit_diagonal` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_lower_triangular(operator) -> bool:
    raise NotImplementedError(f"`is_lower_triangular` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_upper_triangular(operator) -> bool:
    raise NotImplementedError(f"`is_upper_triangular` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_positive_semidefinite(operator) -> bool:
    raise NotImplementedError(f"`is_positive_semidefinite` has not been implemented for {type(operator)}")


@ft.singledispatch
def is_negative_semidefinite(operator) -> bool:
    raise NotImplementedError(f"`is_negative_semidefinite` has not been implemented for {type(operator)}")


_LEAF_OPS = (MatrixLinearOperator, FunctionLinearOperator, JacobianLinearOperator)


def _register_leaf(check, fn):
    for cls in _LEAF_OPS:
        check.register(cls)(fn)


def _sym_leaf(op):
    tags = _tagged(op)
    if symmetric_tag in tags or diagonal_tag in tags:
        return True
    if positive_semidefinite_tag in tags or negative_semidefinite_tag in tags:
        return _has_real_dtype(op)
    return False


_register_leaf(is_symmetric, _sym_leaf)
_register_leaf(is_diagonal, lambda op: diagonal_tag in op.tags or (op.in_size() == 1 and op.out_size() == 1))
_register_leaf(is_tridiagonal, lambda op: tridiagonal_tag in op.tags or diagonal_tag in op.tags)
_register_leaf(has_unit_diagonal, lambda op: unit_diagonal_tag in op.tags)
_register_leaf(is_lower_triangular, lambda op: lower_triangular_tag in op.tags)
_register_leaf(is_upper_triangular, lambda op: upper_triangular_tag in op.tags)
_register_leaf(is_positive_semidefinite, lambda op: positive_semidefinite_tag in op.tags)
_register_leaf(is_negative_semidefinite, lambda op: negative_semidefinite_tag in op.tags)

for _check in (is_symmetric, is_positive_semidefinite):
    _check.register(IdentityLinearOperator)(lambda op: structure_equal(op.in_structure(), op.out_structure()))
for _check in (is_diagonal, is_tridiagonal, is_lower_triangular, is_upper_triangular, has_unit_diagonal):
    _check.register(IdentityLinearOperator)(lambda op: True)
is_negative_semidefinite.register(IdentityLinearOperator)(lambda op: False)

for _check, _tag in (
    (is_symmetric, symmetric_tag),
    (is_diagonal, diagonal_tag),
    (has_unit_diagonal, unit_diagonal_tag),
    (is_lower_triangular, lower_triangular_tag),
    (is_upper_triangular, upper_triangular_tag),
    (is_positive_semidefinite, positive_semidefinite_tag),
    (is_negative_semidefinite, negative_semidefinite_tag),
    (is_tridiagonal, tridiagonal_tag),
):
    _check.register(TaggedLinearOperator)(
        lambda op, _c=_check, _t=_tag: (_t in op.tags) or _c(op.operator)
    )

# Scaling and negation keep structure (lineax/_operator.py:2112-2139).
for _check in (is_symmetric, is_diagonal, is_lower_triangular, is_upper_triangular, is_tridiagonal):
    for _cls in (MulLinearOperator, NegLinearOperator, DivLinearOperator):
        _check.register(_cls)(lambda op, _c=_check: _c(op.operator))
has_unit_diagonal.register(NegLinearOperator)(lambda op: False)


def _scalar_sign(s):
    if isinstance(s, torch.Tensor):
        if s.numel() != 1 or s.requires_grad:
            return None
        s = float(s)
    try:
        s = float(s)
    except (TypeError, ValueError):
        return None
    return 1 if s > 0 else (-1 if s < 0 else 0)


def _unit_scaled(op):
    sign = _scalar_sign(op.scalar)
    return sign is not None and float(op.scalar) == 1.0 and has_unit_diagonal(op.operator)


has_unit_diagonal.register(MulLinearOperator)(_unit_scaled)
has_unit_diagonal.register(DivLinearOperator)(_unit_scaled)


def _psd_scaled(op):
    sign = _scalar_sign(op.scalar)
    if sign == 1:
        return is_positive_semidefinite(op.operator)
    if sign == -1:
        return is_negative_semidefinite(op.operator)
    return False


def _nsd_scaled(op):
    sign = _scalar_sign(op.scalar)
    if sign == 1:
        return is_negative_semidefinite(op.operator)
    if sign == -1:
        return is_positive_semidefinite(op.operator)
    return False


for _cls in (MulLinearOperator, DivLinearOperator):
    is_positive_semidefinite.register(_cls)(_psd_scaled)
    is_negative_semidefinite.register(_cls)(_nsd_scaled)
is_positive_semidefinite.register(NegLinearOperator)(lambda op: is_negative_semidefinite(op.operator))
is_negative_semidefinite.register(NegLinearOperator)(lambda op: is_positive_semidefinite(op.operator))

# Sums keep structure but not a unit diagonal (lineax/_operator.py:2230-2246).
for _check in (
    is_symmetric,
    is_diagonal,
    is_tridiagonal,
    is_lower_triangular,
    is_upper_triangular,
    is_positive_semidefinite,
    is_negative_semidefinite,
):
    _check.register(AddLinearOperator)(lambda op, _c=_check: _c(op.operator1) and _c(op.operator2))
has_unit_diagonal.register(AddLinearOperator)(lambda op: False)

# Composition (lineax/_operator.py:2249-2291).
for _check in (is_diagonal, is_lower_triangular, is_upper_triangular):
    _check.register(ComposedLinearOperator)(lambda op, _c=_check: _c(op.operator1) and _c(op.operator2))
is_symmetric.register(ComposedLinearOperator)(lambda op: is_diagonal(op.operator1) and is_diagonal(op.operator2))


@is_tridiagonal.register(ComposedLinearOperator)
def _(op):
    if is_diagonal(op.operator1):
        return is_tridiagonal(op.operator2)
    if is_diagonal(op.operator2):
        return is_tridiagonal(op.operator1)
    return False


is_positive_semidefinite.register(ComposedLinearOperator)(lambda op: False)
is_negative_semidefinite.register(ComposedLinearOperator)(lambda op: False)
has_unit_diagonal.register(ComposedLinearOperator)(
    lambda op: (is_diagonal(op) or is_lower_triangular(op) or is_upper_triangular(op))
    and has_unit_diagonal(op.operator1)
    and has_unit_diagonal(op.operator2)
)


# ---------------------------------------------------------------------------
# linearise / materialise / conj
# ---------------------------------------------------------------------------


@ft.singledispatch
def linearise(operator):
    """Operator optimised for repeated matvecs (lineax/_operator.py:1206-1257).

    Every operator here is already linear in its representation, so this is
    the identity except for wrappers, which recurse.
    """
    return operator


@linearise.register(TaggedLinearOperator)
def _(operator):
    return TaggedLinearOperator(linearise(operator.operator), operator.tags)


@linearise.register(JacobianLinearOperator)
def _(operator):
    # the primal pass runs once (lineax/_operator.py:1240-1257: jax.linearize,
    # or the transposed vjp for jac="bwd")
    if operator.jac == "bwd":
        y, vjp_fn = torch.func.vjp(operator._f, operator.x)
        if is_symmetric(operator):
            one = lambda v: vjp_fn(v)[0]
        else:
            zeros = tree_map(torch.zeros_like, y)
            one = lambda v: torch.func.vjp(lambda u: vjp_fn(u)[0], zeros)[1](v)[0]
    else:
        with warnings.catch_warnings():
            # torch.func.linearize's graph folding reports an internal detail
            warnings.filterwarnings("ignore", message="Attempted to insert a get_attr Node")
            _, one = torch.func.linearize(operator._f, operator.x)
    return FunctionLinearOperator(_PerVector(one, operator.in_structure()), operator.in_structure(), operator.tags,
                                  output_structure=operator.out_structure(), device=operator.device)


@ft.singledispatch
def materialise(operator):
    """Operator backed by an explicit matrix (lineax/_operator.py:1263-1315)."""
    if isinstance(operator, (MatrixLinearOperator, IdentityLinearOperator)):
        return operator
    return MatrixLinearOperator(operator.as_matrix(), _tagged(operator))


@materialise.register(TaggedLinearOperator)
def _(operator):
    return TaggedLinearOperator(materialise(operator.operator), operator.tags)


@materialise.register(JacobianLinearOperator)
def _(operator):
    # lineax/_operator.py:1353-1366; a single 1-D input and output gives a
    # MatrixLinearOperator, other structures a FunctionLinearOperator over
    # the flattened matrix
    jac = _jacobian_matrix(operator)
    ins, outs = operator.in_structure(), operator.out_structure()
    if isinstance(ins, TensorStruct) and isinstance(outs, TensorStruct) and len(ins.shape) == 1 \
            and len(outs.shape) == 1:
        return MatrixLinearOperator(jac, operator.tags)

    def apply(v):
        batch = batch_shape_of(v, ins)
        return unravel(matvec(jac, ravel(v, ins, batch)), outs, batch)

    return FunctionLinearOperator(apply, ins, operator.tags, output_structure=outs, device=operator.device)


def conj(operator):
    """Complex conjugate of an operator; the identity for real dtypes.

    lineax/_operator.py:2301-2408.  Complex dtypes are not supported by the
    kernels of this package.
    """
    if _has_real_dtype(operator):
        return operator
    raise NotImplementedError("complex operators are not supported by lineax_amd kernels")


def diagonal(operator) -> torch.Tensor:
    """Diagonal of the operator's matrix."""
    return torch.diagonal(operator.as_matrix(), dim1=-2, dim2=-1)


def mv_into(operator, v: torch.Tensor, struct, out: torch.Tensor) -> torch.Tensor:
    """operator.mv on a flat [rows, n] tensor, written into ``out`` (flat,
    contiguous, in the storage dtype).  Operators with an ``mv_into`` method
    (the stencil operators) write there directly; others are copied in."""
    if isinstance(operator, TaggedLinearOperator):
        return mv_into(operator.operator, v, struct, out)
    direct = getattr(operator, "mv_into", None)
    if direct is not None and isinstance(struct, TensorStruct):
        direct(v, out)
        return out
    if isinstance(operator, NegLinearOperator):
        return mv_into(operator.operator, v, struct, out).neg_()
    return out.copy_(mv_as(operator, v, struct, out.dtype))


def mv_as(operator, v: torch.Tensor, struct, out_dtype: torch.dtype) -> torch.Tensor:
    """operator.mv on a flat [rows, n] tensor, returned flat in ``out_dtype``.

    Operators with an ``mv_as`` method (the stencil operators) compute in
    the requested storage dtype directly; negation and tag wrappers pass it
    through; any other operator is called on its own dtype with casts.
    """
    rows = v.shape[0]
    if isinstance(operator, TaggedLinearOperator):
        return mv_as(operator.operator, v, struct, out_dtype)
    if isinstance(operator, NegLinearOperator):
        return -mv_as(operator.operator, v, struct, out_dtype)
    direct = getattr(operator, "mv_as", None)
    if direct is not None and isinstance(struct, TensorStruct):
        return direct(v.reshape(rows, *struct.shape), out_dtype).reshape(rows, -1)
    out = operator.mv(unravel(v, struct, (rows,)))
    return ravel(out, operator.out_structure(), (rows,)).to(out_dtype)
