"""Large dense LU: butterfly-randomised, window-pivoted, refined.

``LargeLU(a)`` factors one float32 (or float64, factored in float32 and
refined against the float64 operator) matrix of order n (hundreds to
thousands):

1. A' = U^T A V with two-sided depth-2 random butterflies (seeded), padded
   to a multiple of 4 with an identity block;
2. P A' = L U, each 128-column block column pivoting over its diagonal
   block and the 128 rows below it (every multiplier bounded by 1),
   blocked: float32 panels and triangular block products, Schur-complement
   products on the matrix cores with float32 accumulation in bfloat16,
   float16, split bfloat16 ("bf16x3": hi*hi + hi*lo + lo*hi, the default) or
   float32 (``precision``).  Plain bf16 / fp16 factors are too inexact for
   ill-conditioned operators (GMRES then needs hundreds of steps, and fp16
   can overflow); bf16x3 keeps float32-level factors at bf16 MFMA rates;
3. ``solve`` refines against the original A: FP64 residuals r = b - A x
   (float32 A, FP64 accumulation), each correction a plain refinement
   step x += V (L U)^-1 P U^T r while those divide the backward error by
   at least 10 (bf16x3 and float32 factors), then from GMRES on A
   preconditioned by the same map
   (float32 basis, FP64 inner products), until the normwise backward error
   eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol or ``max_refine``
   residuals.

The answer is that of lineax's LU (lineax/_solver/lu.py: getrf with partial
pivoting, then getrs) to the refinement tolerance; the factors are not
getrf's.  Pivoting over a 256-row window of the randomised matrix replaces
one device-wide arg-max per column with a workgroup-local one.
"""

from __future__ import annotations

import math

import torch

from . import _kernels as K
from ._operator import FunctionLinearOperator
from ._structure import TensorStruct

MAX_RHS_PER_SOLVE = 8


class _F32Matrix(FunctionLinearOperator):
    """x -> A x for a float32 matrix, in the dtype asked for (FP64 accumulation)."""

    def __init__(self, a: torch.Tensor, dtype: torch.dtype):
        self.a = a
        n = a.shape[0]
        struct = TensorStruct((n,), dtype)
        super().__init__(self._apply, struct, (), output_structure=struct, device=a.device)

    def _apply(self, x):
        return self.mv_as(x.reshape(-1, self.a.shape[1]), x.dtype).reshape(x.shape)

    def mv_as(self, x: torch.Tensor, out_dtype: torch.dtype) -> torch.Tensor:
        return K.gemv_f32a(self.a, x.reshape(-1, self.a.shape[1]), out_dtype).reshape(*x.shape[:-1], self.a.shape[0])


class _F64Matrix(_F32Matrix):
    """x -> A x for a float64 matrix (lxa_gemv in FP64), in the dtype asked for."""

    def mv_as(self, x: torch.Tensor, out_dtype: torch.dtype) -> torch.Tensor:
        y = K.gemv(self.a, x.reshape(-1, self.a.shape[1]).to(torch.float64))
        return y.to(out_dtype).reshape(*x.shape[:-1], self.a.shape[0])


class _Preconditioner(FunctionLinearOperator):
    """x -> V (L U)^-1 U^T x (or its transpose), float32 arithmetic."""

    def __init__(self, fac: "LargeLU", dtype: torch.dtype, trans: bool):
        self.fac = fac
        self.trans = trans
        struct = TensorStruct((fac.n,), dtype)
        super().__init__(self._apply, struct, (), output_structure=struct, device=fac.lu.device)

    def _apply(self, x):
        return self.mv_as(x.reshape(-1, self.fac.n), x.dtype).reshape(x.shape)

    def mv_as(self, x: torch.Tensor, out_dtype: torch.dtype) -> torch.Tensor:
        y = self.fac.apply_inverse(x.reshape(-1, self.fac.n).to(torch.float32), trans=self.trans)
        return y.to(out_dtype).reshape(*x.shape[:-1], self.fac.n)


class LargeLU:
    """Factorisation of one large float32 matrix (see the module docstring)."""

    def __init__(self, a: torch.Tensor, precision: str = "bf16x3", seed: int = 0):
        K._gpu(a, "matrix")
        if a.dim() != 2 or a.shape[0] != a.shape[1]:
            raise ValueError(f"LargeLU needs one square matrix, got {tuple(a.shape)}")
        if a.dtype not in (torch.float32, torch.float64):
            raise TypeError("LargeLU factors float32 or float64 matrices")
        if precision not in K.GEMM_MODES:
            raise ValueError(f"precision must be one of {sorted(K.GEMM_MODES)}")
        self.precision = precision
        self.a = a.detach().contiguous()
        self.n = n = a.shape[0]
        self.np = np_ = -(-n // 4) * 4
        dev = a.device
        # a float64 operator is factored in float32 and refined against
        # itself (FP64 residuals and products)
        lu = torch.zeros((np_, np_), dtype=torch.float32, device=dev)
        lu[:n, :n] = self.a
        if np_ > n:
            idx = torch.arange(n, np_, device=dev)
            lu[idx, idx] = 1.0
        gen = torch.Generator(device="cpu").manual_seed(int(seed))
        # butterfly diagonals exp(r / 10), r ~ U(-1/2, 1/2) (Parker 1995)
        self.du = torch.exp((torch.rand(2 * np_, generator=gen, dtype=torch.float64) - 0.5) / 10).float().to(dev)
        self.dv = torch.exp((torch.rand(2 * np_, generator=gen, dtype=torch.float64) - 0.5) / 10).float().to(dev)
        K.rbt_matrix(lu, self.du, self.dv)
        blk = K.lu_large_block()
        self.nblk = -(-np_ // blk)
        self.linv = torch.empty(self.nblk * blk * blk, dtype=torch.float32, device=dev)
        self.uinv = torch.empty_like(self.linv)
        self.info = torch.zeros(1, dtype=torch.int32, device=dev)
        self.piv = torch.empty(np_, dtype=torch.int32, device=dev)
        K.lu_large_factor(lu, precision, self.linv, self.uinv, self.piv, self.info)
        self.perm = K.lu_large_perm(self.piv)
        self.lu = lu
        self.sync = K.lu_large_trsv_workspace(np_, dev)
        # set on the device when a block triangular solve gave up waiting for a
        # dependency (the kernel's bounded spin); checked once per solve
        self.trsv_failed = torch.zeros(1, dtype=torch.int32, device=dev)
        self.norm_inf = self.a.abs().sum(dim=1, dtype=torch.float64).max()

    def apply_inverse(self, v: torch.Tensor, trans: bool = False) -> torch.Tensor:
        """A^-1 v through the factors: V (LU)^-1 P U^T v, or U P^T (LU)^-T V^T v."""
        rows = v.shape[0]
        w = torch.zeros((rows, self.np), dtype=torch.float32, device=v.device)
        w[:, : self.n] = v
        first, last = (self.dv, self.du) if trans else (self.du, self.dv)
        K.rbt_vectors(w, first, form=1)
        if not trans:
            w = K.lu_large_permute(w, self.perm, inverse=False)
        for r0 in range(0, rows, MAX_RHS_PER_SOLVE):
            chunk = w[r0: r0 + MAX_RHS_PER_SOLVE]
            for kind, inv in (((2, self.uinv), (3, self.linv)) if trans else ((0, self.linv), (1, self.uinv))):
                K.lu_large_trsv(kind, self.lu, inv, chunk, self.sync)
                fw = K.LU_TRSV_FAILED_WORD
                torch.maximum(self.trsv_failed, self.sync[fw: fw + 1], out=self.trsv_failed)
        if trans:
            w = K.lu_large_permute(w, self.perm, inverse=True)
        K.rbt_vectors(w, last, form=0)
        return w[:, : self.n]

    def residual(self, x: torch.Tensor, b: torch.Tensor, trans: bool = False) -> torch.Tensor:
        """b - A x (or b - A^T x) in FP64."""
        if self.a.dtype == torch.float64:
            return b.reshape(-1, self.n) - K.gemv(self.a, x.reshape(-1, self.n), trans=trans)
        a = self.a.T.contiguous() if trans else self.a
        return K.gemv_f32a(a, x, torch.float64, b=b)

    def backward_error(self, x: torch.Tensor, b: torch.Tensor, r: torch.Tensor, trans: bool = False):
        norm_a = self.a.abs().sum(dim=0, dtype=torch.float64).max() if trans else self.norm_inf
        num = r.abs().amax(dim=1)
        den = norm_a * x.abs().amax(dim=1) + b.abs().amax(dim=1)
        return torch.where(num == 0, torch.zeros_like(num), num / den)

    def solve(self, b: torch.Tensor, *, tol: float = 1e-12, max_refine: int = 10, trans: bool = False,
              inner_rtol: float = 1e-4, restart: int = 4, inner_max_steps: int = 50, plain_first: bool | None = None,
              plain_gain: float = 0.1):
        """Refined solution x (float64, [nrhs, n]) of A x = b (A^T x = b if trans).

        With plain_first (the default for bf16x3 and float32 factors),
        corrections start as plain iterative refinement, x += (LU)^-1 r: one
        preconditioner application each (bf16x3 factors of Gaussian
        operators of order 8192 divide the backward error by about 10 per
        step; bf16 and fp16 factors by less than 2, so they start with
        GMRES).  Once a plain step divides it by less than 1 / plain_gain on
        a right-hand side that has not converged (or plain_first is off),
        the remaining corrections are GMRES-IR (Carson &
        Higham 2017): GMRES in float32 on the preconditioned FP64 residual
        equation, restart cycles of ``restart`` Arnoldi iterations (with
        factors this close to A a few iterations per correction suffice, so
        short cycles avoid preconditioner applications that lineax's
        cycle-end convergence test would otherwise spend).  A right-hand side
        whose last plain step raised its backward error goes back to the
        iterate before that step.

        Returns (x, eta, stats) with eta the last backward error per
        right-hand side; stats counts residuals, plain steps and GMRES steps.
        """
        from . import _solve, _solver

        b64 = b.reshape(-1, self.n).to(torch.float64).contiguous()
        x = self.apply_inverse(b64.to(torch.float32), trans=trans).to(torch.float64)
        mat = _F64Matrix if self.a.dtype == torch.float64 else _F32Matrix
        op = mat(self.a.T.contiguous() if trans else self.a, torch.float64)
        prec = _Preconditioner(self, torch.float64, trans)
        residuals = 0
        plain_steps = 0
        gmres_steps = 0
        eta = None
        plain = self.precision in ("bf16x3", "fp32") if plain_first is None else bool(plain_first)
        before = None  # (x, r, eta) before the last plain step
        for _ in range(max(1, max_refine)):
            r = self.residual(x, b64, trans)
            residuals += 1
            eta = self.backward_error(x, b64, r, trans)
            if before is not None:
                x0, r0, eta0 = before
                before = None
                if not bool(((eta <= plain_gain * eta0) | (eta <= tol)).all()):
                    plain = False
                    worse = ~(eta <= eta0)  # NaN counts as worse
                    x = torch.where(worse[:, None], x0, x)
                    r = torch.where(worse[:, None], r0, r)
                    eta = torch.where(worse, eta0, eta)
            # stop when converged, out of residuals, or non-finite (a zero
            # pivot: the output stays non-finite, reported as singular)
            if bool((eta <= tol).all()) or residuals >= max_refine or not bool(torch.isfinite(eta).all()):
                break
            if plain:
                before = (x, r, eta)
                x = x + prec.mv_as(r, torch.float64)
                plain_steps += 1
                continue
            d0 = prec.mv_as(r, torch.float64)
            atol = inner_rtol * float(d0.abs().max())
            sol = _solve.linear_solve(
                op, r, _solver.GMRES(rtol=inner_rtol, atol=atol, restart=restart, max_steps=inner_max_steps,
                                     working_dtype=torch.float32),
                options={"preconditioner": prec, "y0": d0}, throw=False,
            )
            gmres_steps += int(sol.stats["num_steps"].max())
            x = x + sol.value
        self.check_trsv()
        return x, eta, {"residuals": residuals, "plain_steps": plain_steps, "gmres_steps": gmres_steps,
                        "zero_pivot": int(self.info.item())}

    def check_trsv(self) -> None:
        """Raises if any block triangular solve since construction timed out
        waiting for a dependency (its output would be wrong); synchronises."""
        if int(self.trsv_failed.item()):
            raise RuntimeError("lxa_lu_large_trsv: a block waited past its spin limit for the block it depends "
                               "on; the triangular solve is incomplete")


def factor_flops(n: int) -> float:
    """Algorithmic flops of an LU of order n with its solves of 4 right-hand sides (2n^3/3 + 8n^2)."""
    return 2.0 * n ** 3 / 3.0 + 8.0 * n ** 2


def padded_order(n: int) -> int:
    return int(math.ceil(n / 4) * 4)
