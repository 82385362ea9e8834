"""Shape/dtype structures and pytree plumbing.

``TensorStruct`` plays the part of ``jax.ShapeDtypeStruct``.  Vectors may carry
leading batch dimensions in front of the structure's shape: that is how a
batch of independent solves is expressed (lineax expresses it with
``jax.vmap``; here one call over the batch launches one batched kernel).
"""

from __future__ import annotations

import dataclasses
import math
from typing import Any

import torch
import torch.utils._pytree as pytree


@dataclasses.dataclass(frozen=True)
class TensorStruct:
    shape: tuple
    dtype: torch.dtype

    def __post_init__(self):
        object.__setattr__(self, "shape", tuple(int(s) for s in self.shape))

    @property
    def size(self) -> int:
        return math.prod(self.shape)


def default_floating_dtype() -> torch.dtype:
    """torch's default float dtype (lineax/_misc.py:73-77 uses JAX's x64 flag)."""
    return torch.get_default_dtype()


def inexact_asarray(x) -> torch.Tensor:
    """Tensor with an inexact dtype; integers become the default float (lineax/_misc.py:80-84)."""
    if not isinstance(x, torch.Tensor):
        x = torch.as_tensor(x)
    if not (x.is_floating_point() or x.is_complex()):
        x = x.to(default_floating_dtype())
    return x


def tree_flatten(tree):
    return pytree.tree_flatten(tree)


def tree_unflatten(leaves, spec):
    return pytree.tree_unflatten(leaves, spec)


def tree_leaves(tree) -> list:
    return pytree.tree_leaves(tree)


def tree_map(fn, tree, *rest):
    return pytree.tree_map(fn, tree, *rest)


def structure_of(tree, batch_ndim: int = 0):
    """Pytree of TensorStruct of ``tree``'s leaves, dropping ``batch_ndim`` leading dims."""
    return tree_map(lambda x: TensorStruct(tuple(x.shape[batch_ndim:]), x.dtype), tree)


def structure_equal(a, b) -> bool:
    la, sa = tree_flatten(a)
    lb, sb = tree_flatten(b)
    return sa == sb and all(x == y for x, y in zip(la, lb))


def batch_shape_of(tree, struct) -> tuple:
    """Leading batch shape of ``tree`` relative to structure ``struct``.

    Raises ValueError if the leaves do not match the structure's trailing
    shapes and dtypes, or disagree on the batch shape.
    """
    leaves, spec = tree_flatten(tree)
    sleaves, sspec = tree_flatten(struct)
    if spec != sspec or len(leaves) != len(sleaves):
        raise ValueError(f"pytree structure {spec} does not match {sspec}")
    batch = None
    for x, s in zip(leaves, sleaves):
        k = len(s.shape)
        if x.dim() < k or tuple(x.shape[x.dim() - k:]) != s.shape or x.dtype != s.dtype:
            raise ValueError(
                f"Vector and operator structures do not match. Got a leaf of shape "
                f"{tuple(x.shape)} and dtype {x.dtype}; expected trailing shape {s.shape} "
                f"and dtype {s.dtype}"
            )
        b = tuple(x.shape[: x.dim() - k])
        if batch is None:
            batch = b
        elif b != batch:
            raise ValueError(f"inconsistent batch shapes {batch} and {b} across vector leaves")
    return batch if batch is not None else ()


def ravel(tree, struct, batch: tuple) -> torch.Tensor:
    """Concatenates the leaves of ``tree`` into [*batch, size] in their common dtype."""
    leaves = tree_leaves(tree)
    if len(leaves) == 1:
        x = leaves[0]
        return x.reshape(*batch, -1) if x.dim() else x.reshape(1)
    dtype = leaves[0].dtype
    for x in leaves[1:]:
        dtype = torch.promote_types(dtype, x.dtype)
    return torch.cat([x.to(dtype).reshape(*batch, -1) for x in leaves], dim=-1)


def unravel(flat: torch.Tensor, struct, batch: tuple):
    """Inverse of ``ravel`` onto ``struct`` (casting each piece to its leaf dtype)."""
    sleaves, spec = tree_flatten(struct)
    sizes = [s.size for s in sleaves]
    pieces = torch.split(flat, sizes, dim=-1) if len(sizes) > 1 else (flat,)
    out = [p.reshape(*batch, *s.shape).to(s.dtype) for p, s in zip(pieces, sleaves)]
    return tree_unflatten(out, spec)


def zeros_like_struct(struct, batch: tuple, device) -> Any:
    return tree_map(lambda s: torch.zeros((*batch, *s.shape), dtype=s.dtype, device=device), struct)


def full_like_tree(tree, value):
    return tree_map(lambda x: torch.full_like(x, value), tree)
