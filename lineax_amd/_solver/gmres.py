"""GMRES solver (lineax/_solver/gmres.py) driving the device Krylov kernels."""

from __future__ import annotations

import torch

from .. import _kernels as K
from .._norm import max_norm
from .._operator import conj, linearise
from .._solve import AbstractLinearSolver
from .._structure import batch_shape_of, ravel, structure_equal, unravel
from .cg import FlatOperator, norm_kind, stencil_of
from .misc import preconditioner_and_y0

MAX_RESTART = 128
# fused stencil Arnoldi steps: restart and grid-width limits
FUSED_MAX_RESTART = 32
FUSED_MAX_WIDTH = 1024


class GMRES(AbstractLinearSolver):
    """Restarted GMRES with lineax's rules (lineax/_solver/gmres.py:39-447).

    Each restart cycle builds an Arnoldi basis by single-pass classical
    Gram-Schmidt, solves the Hessenberg least-squares problem by Householder
    QR, and updates the iterate.  The first cycle only forms the residual;
    breakdown ends a cycle early and is reported one cycle later unless the
    tolerance is met; ``stagnation_iters`` cycles without a decrease of the
    residual norm stop the solve.  ``working_dtype`` (not in lineax) stores
    the basis and residual in a narrower dtype than the vector's; inner
    products, the Hessenberg matrix and the least-squares solve are FP64.
    ``graph`` (not in lineax) captures a restart cycle in a HIP graph and
    replays it, checking the active flags on the host between cycles.
    ``fused`` (not in lineax): for a five-point stencil operator with no
    preconditioner (grid width <= 1024, restart <= 32), each Arnoldi step is
    one sweep of lxa_gmres_stencil_step that reads the basis once.
    """

    def __init__(self, rtol, atol, norm=max_norm, max_steps: int | None = None, restart: int = 20,
                 stagnation_iters: int = 20, working_dtype: torch.dtype | None = None, graph: bool = False,
                 fused: bool = True):
        if isinstance(rtol, (int, float)) and rtol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if isinstance(atol, (int, float)) and atol < 0:
            raise ValueError("Tolerances must be non-negative.")
        if atol == 0 and rtol == 0 and max_steps is None:
            raise ValueError("Must specify `rtol`, `atol`, or `max_steps` (or some combination of all three).")
        self.rtol = rtol
        self.atol = atol
        self.norm = norm
        self.max_steps = max_steps
        self.restart = restart
        self.stagnation_iters = stagnation_iters
        self.working_dtype = working_dtype
        self.graph = graph
        self.fused = fused

    def init(self, operator, options):
        if not structure_equal(operator.in_structure(), operator.out_structure()):
            raise ValueError("`GMRES(..., normal=False)` may only be used for linear solves with square matrices.")
        return linearise(operator)

    def compute(self, state, vector, options):
        operator = state
        preconditioner, y0 = preconditioner_and_y0(operator, vector, options)
        struct = operator.in_structure()
        batch = batch_shape_of(vector, struct)
        b = ravel(vector, struct, batch)
        size = b.shape[-1]
        rows = b.numel() // size if size else 0
        b = b.reshape(rows, size).contiguous()
        ydt = b.dtype
        vdt = self.working_dtype or ydt
        dev = b.device
        max_steps = 10 * size if self.max_steps is None else int(self.max_steps)
        m = min(int(self.restart), size)
        if m > MAX_RESTART:
            raise NotImplementedError(f"GMRES restart is limited to {MAX_RESTART} here (got {m})")
        prm = K.KrylovParams(
            rtol=float(self.rtol), atol=float(self.atol), rcond=0.0, eps=float(torch.finfo(vdt).eps),
            max_steps=max_steps, max_steps_given=int(self.max_steps is not None), norm_kind=norm_kind(self.norm),
            restart=m, stagnation_iters=int(self.stagnation_iters),
        )
        A = FlatOperator(operator, struct, batch)
        M = FlatOperator(preconditioner, struct, batch)
        precond = not M.identity
        yd, vd = K.dtype_code(ydt), K.dtype_code(vdt)
        st = K.RowState(rows, dev)
        gs = K.RowState(rows, dev)
        part = K.partials_buffer(rows, size, m, dev)
        stream = K.current_stream(b)
        basis = torch.empty((rows, m + 1, size), dtype=vdt, device=dev)
        hess = torch.empty((rows, m, m + 1), dtype=torch.float64, device=dev)
        proj = torch.empty((rows, m + 1), dtype=torch.float64, device=dev)
        z = torch.empty((rows, m), dtype=torch.float64, device=dev)
        r = torch.empty((rows, size), dtype=vdt, device=dev)
        y = ravel(y0, struct, batch).reshape(rows, size).to(ydt).contiguous().clone()

        K.call("lxa_gmres_init", stream, rows, st, gs)

        def residual():
            stream = K.current_stream(b)
            t = A(y, ydt)
            K.call("lxa_gmres_residual", stream, yd, vd, rows, size, b, t, r, 1, prm, st, part)
            if precond:
                r.copy_(M(r, vdt))
                K.call("lxa_gmres_residual", stream, yd, vd, rows, size, b, t, r, 0, prm, st, part)

        stencil = None
        if self.fused and not precond and dev.type == "cuda" and m <= FUSED_MAX_RESTART:
            stencil = stencil_of(operator)
            if stencil is not None and (stencil[0] * stencil[1] != size or stencil[1] > FUSED_MAX_WIDTH):
                stencil = None

        if stencil is None:
            def cycle():
                stream = K.current_stream(b)  # the capture stream while a graph is recorded
                K.call("lxa_gmres_cycle_begin", stream, vd, rows, size, r, basis, prm, st, gs, hess, part)
                for j in range(m):
                    w = M(A(basis[:, j, :], vdt), vdt).contiguous()
                    K.call("lxa_gmres_arnoldi", stream, vd, rows, size, j, basis, w, prm, st, gs, hess, proj, part)
                K.call("lxa_gmres_update", stream, yd, vd, rows, size, basis, y, prm, st, gs, hess, z, part)
                residual()
                K.call("lxa_gmres_cycle_end", stream, rows, size, 0, prm, st, gs, part)
        else:
            nx, ny, coeffs = stencil
            ws = K.StencilGMRESWorkspace(rows, nx, ny, m, coeffs, dev)
            wbuf = (torch.empty((rows, size), dtype=vdt, device=dev), torch.empty((rows, size), dtype=vdt, device=dev))

            def cycle():
                stream = K.current_stream(b)
                K.call("lxa_gmres_cycle_begin", stream, vd, rows, size, r, basis, prm, st, gs, hess, part)
                for j in range(m + 1):
                    mode = 0 if j == 0 else (2 if j == m else 1)
                    w_in = wbuf[(j - 1) % 2] if j > 0 else None
                    w_out = wbuf[j % 2] if j < m else None
                    K.call("lxa_gmres_stencil_step", stream, vd, rows, nx, ny, ws.coeffs, mode, j, prm, basis, w_in,
                           w_out, ws.coefs, ws.colscale, proj, hess, st, gs, ws.partials)
                K.call("lxa_gmres_update_scaled", stream, yd, vd, rows, size, basis, y, prm, st, gs, hess, z,
                       ws.colscale, part)
                residual()
                K.call("lxa_gmres_cycle_end", stream, rows, size, 0, prm, st, gs, part)

        # The dummy first pass only forms r0 (gmres.py:313-317).
        residual()
        K.call("lxa_gmres_cycle_end", stream, rows, size, 1, prm, st, gs, part)
        cycles = 1
        graph = None
        use_graph = self.graph and dev.type == "cuda"
        while cycles < max_steps and bool(st.active().any()):
            if graph is not None:
                graph.replay()
            elif use_graph and cycles >= 2:
                # one restart cycle captured once (after an eager one), then replayed
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    cycle()
                graph.replay()
            else:
                cycle()
            cycles += 1
        stream = K.current_stream(b)
        K.call("lxa_gmres_result", stream, rows, prm, st, gs)

        solution = unravel(y.reshape(*batch, size), struct, batch)
        result = st.result().clone().reshape(batch)
        stats = {"num_steps": st.step().clone().reshape(batch), "max_steps": self.max_steps}
        return solution, result, stats

    def transpose(self, state, options):
        transpose_options = {}
        if "preconditioner" in options:
            transpose_options["preconditioner"] = options["preconditioner"].transpose()
        return state.transpose(), transpose_options

    def conj(self, state, options):
        conj_options = {}
        if "preconditioner" in options:
            conj_options["preconditioner"] = conj(options["preconditioner"])
        return conj(state), conj_options

    def assume_full_rank(self):
        return True
