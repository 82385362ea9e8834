"""Solver helpers (lineax/_solver/misc.py:30-114).

Structures are packed as ``(out_structure, in_structure)``; vectors are
flattened to ``[*batch, size]`` for the kernels and restored afterwards.
"""

from __future__ import annotations

from .._operator import AbstractLinearOperator, IdentityLinearOperator, linearise
from .._structure import (
    batch_shape_of,
    ravel,
    structure_equal,
    tree_map,
    unravel,
)

import torch


def preconditioner_and_y0(operator: AbstractLinearOperator, vector, options: dict):
    """Preconditioner (default identity) and initial guess (default zeros).

    lineax/_solver/misc.py:30-60.
    """
    structure = operator.in_structure()
    if "preconditioner" in options:
        preconditioner = linearise(options["preconditioner"])
        if not isinstance(preconditioner, AbstractLinearOperator):
            raise ValueError("The preconditioner must be a linear operator.")
        if not structure_equal(preconditioner.in_structure(), structure):
            raise ValueError(
                "The preconditioner must have `in_structure` that matches the operator's `in_strucure`."
            )
        if not structure_equal(preconditioner.out_structure(), structure):
            raise ValueError(
                "The preconditioner must have `out_structure` that matches the operator's `in_structure`."
            )
    else:
        preconditioner = IdentityLinearOperator(structure)
    if "y0" in options:
        y0 = options["y0"]
        if not structure_equal(
            tree_map(lambda t: (tuple(t.shape), t.dtype), y0), tree_map(lambda t: (tuple(t.shape), t.dtype), vector)
        ):
            raise ValueError("`y0` must have the same structure, shape, and dtype as `vector`")
    else:
        y0 = tree_map(torch.zeros_like, vector)
    return preconditioner, y0


def pack_structures(operator: AbstractLinearOperator):
    return (operator.out_structure(), operator.in_structure())


def transpose_packed_structures(packed):
    out_structure, in_structure = packed
    return (in_structure, out_structure)


def ravel_vector(vector, packed) -> tuple[torch.Tensor, tuple]:
    """[*batch, out_size] in the vector's common dtype, and the batch shape."""
    out_structure, _ = packed
    try:
        batch = batch_shape_of(vector, out_structure)
    except ValueError as e:
        raise ValueError("pytree does not match out_structure") from e
    return ravel(vector, out_structure, batch), batch


def unravel_solution(solution: torch.Tensor, packed, batch: tuple):
    _, in_structure = packed
    return unravel(solution, in_structure, batch)
