"""Norms and inner products over pytrees of tensors (lineax/_norm.py).

Reductions accumulate in float64.  These are the general-purpose helpers of
the public API; the Krylov solvers use fused kernels for their per-iteration
reductions.
"""

from __future__ import annotations

import math

import torch

from ._structure import default_floating_dtype, tree_flatten, tree_leaves


def tree_dot(tree1, tree2) -> torch.Tensor:
    """sum(conj(x) * y) over all leaves (lineax/_norm.py:27-47)."""
    l1, s1 = tree_flatten(tree1)
    l2, s2 = tree_flatten(tree2)
    if s1 != s2:
        raise ValueError("trees must have the same structure")
    if not l1:
        return torch.tensor(0.0, dtype=default_floating_dtype())
    out = None
    for a, b in zip(l1, l2):
        d = torch.sum(torch.conj(a).reshape(-1).to(torch.float64) * b.reshape(-1).to(torch.float64))
        out = d if out is None else out + d
    dtype = torch.promote_types(l1[0].dtype, l2[0].dtype)
    return out.to(dtype)


def sum_squares(x) -> torch.Tensor:
    return tree_dot(x, x).real


def two_norm(x) -> torch.Tensor:
    """L2 norm of all leaves (lineax/_norm.py:59-101)."""
    leaves = tree_leaves(x)
    size = sum(t.numel() for t in leaves)
    if size == 1:
        for t in leaves:
            if t.numel() == 1:
                return torch.abs(t.reshape(()))
    return torch.sqrt(sum_squares(x))


def rms_norm(x) -> torch.Tensor:
    """two_norm / sqrt(size) (lineax/_norm.py:104-120)."""
    leaves = tree_leaves(x)
    size = sum(t.numel() for t in leaves)
    if size == 0:
        dtype = default_floating_dtype() if not leaves else leaves[0].dtype
        return torch.tensor(0.0, dtype=dtype)
    return two_norm(x) / math.sqrt(size)


def max_norm(x) -> torch.Tensor:
    """Largest absolute entry over all leaves (lineax/_norm.py:123-139)."""
    leaves = [t for t in tree_leaves(x) if t.numel() > 0]
    if not leaves:
        all_leaves = tree_leaves(x)
        dtype = default_floating_dtype() if not all_leaves else all_leaves[0].dtype
        return torch.tensor(0.0, dtype=dtype)
    out = None
    for t in leaves:
        m = torch.max(torch.abs(t))
        out = m if out is None else torch.maximum(out, m)
    return out
