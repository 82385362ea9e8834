"""Matrix-free grid operators for the Krylov workloads.

``StencilOperator`` is a ``FunctionLinearOperator`` whose function is the
five-point stencil kernel: on an nx x ny grid (j the fast index) with zero
Dirichlet boundary,

    (A u)[i, j] = c u[i, j] + w u[i, j-1] + e u[i, j+1] + s u[i-1, j] + n u[i+1, j].

Vectors are grid functions of shape (nx, ny), with optional leading batch
dimensions (one grid per right-hand side).

``poisson2d``: c = 4, w = e = s = n = -1 (the unscaled 5-point Laplacian,
symmetric positive definite).

``convection_diffusion2d``: -(1/Pe) Laplacian + v . grad, first-order upwind,
on the unit square with h = 1/(n+1), multiplied through by h^2:
    c = 4/Pe + (|vx| + |vy|) h,  neighbours -1/Pe, and the upwind neighbour
    of each direction gets an extra -|v| h (j-1 for vx > 0, i-1 for vy > 0).
"""

from __future__ import annotations

import torch

from . import _kernels as K
from ._operator import FunctionLinearOperator, _materialise_by_columns
from ._structure import TensorStruct
from ._tags import as_tag_set, positive_semidefinite_tag, symmetric_tag, transpose_tags


class StencilOperator(FunctionLinearOperator):
    """Five-point stencil operator with constant coefficients."""

    def __init__(self, nx: int, ny: int, coeffs, *, dtype=torch.float64, device="cuda", tags=()):
        self.nx, self.ny = int(nx), int(ny)
        self.coeffs = tuple(float(c) for c in coeffs)
        if len(self.coeffs) != 5:
            raise ValueError("coeffs = (center, west, east, south, north)")
        struct = TensorStruct((self.nx, self.ny), dtype)
        super().__init__(self._apply, struct, as_tag_set(tags), output_structure=struct, device=torch.device(device))

    def _apply(self, x, out_dtype=None):
        return K.stencil5(x, self.nx, self.ny, self.coeffs, out_dtype=out_dtype)

    def mv_as(self, x: torch.Tensor, out_dtype: torch.dtype) -> torch.Tensor:
        """The product in another storage dtype (FP64 accumulation either way)."""
        return self._apply(x, out_dtype=out_dtype)

    def mv_into(self, x: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """The product written into ``out`` (its dtype is the storage dtype)."""
        return K.stencil5(x, self.nx, self.ny, self.coeffs, out=out)

    def as_matrix(self):
        return _materialise_by_columns(self)

    def transpose(self):
        c, w, e, s, n = self.coeffs
        if w == e and s == n:
            return self
        struct = self.in_structure()
        return StencilOperator(self.nx, self.ny, (c, e, w, n, s), dtype=struct.dtype, device=self.device,
                               tags=transpose_tags(self.tags))

    def __repr__(self):
        return f"StencilOperator({self.nx}x{self.ny}, coeffs={self.coeffs}, tags={set(self.tags)})"


def poisson2d(nx: int, ny: int | None = None, *, dtype=torch.float64, device="cuda") -> StencilOperator:
    ny = nx if ny is None else ny
    return StencilOperator(nx, ny, (4.0, -1.0, -1.0, -1.0, -1.0), dtype=dtype, device=device,
                           tags=(positive_semidefinite_tag, symmetric_tag))


def convection_diffusion_coeffs(n: int, peclet: float = 10.0, velocity=(1.0, 1.0)):
    """(center, west, east, south, north) of the upwind operator scaled by h^2."""
    h = 1.0 / (n + 1)
    vx, vy = float(velocity[0]), float(velocity[1])
    nu = 1.0 / peclet
    c = 4 * nu + (abs(vx) + abs(vy)) * h
    w = e = s = nn = -nu
    if vx > 0:
        w -= vx * h
    elif vx < 0:
        e += vx * h
    if vy > 0:
        s -= vy * h
    elif vy < 0:
        nn += vy * h
    return (c, w, e, s, nn)


def convection_diffusion2d(n: int, *, peclet: float = 10.0, velocity=(1.0, 1.0), dtype=torch.float64,
                           device="cuda") -> StencilOperator:
    return StencilOperator(n, n, convection_diffusion_coeffs(n, peclet, velocity), dtype=dtype, device=device)
