"""ctypes boundary to the gfx950 kernel library (``include/lineax_amd.h``).

Every wrapper checks devices, dtypes, shapes and strides on the host before
a launch, passes torch's current HIP stream, and raises ``RuntimeError`` with
the library's message on a non-zero status.  There is no CPU fallback: a
missing library or a CPU tensor is an error.
"""

from __future__ import annotations

import ctypes
import math
import pathlib
import threading

import numpy as np
import torch

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "liblineax_amd.so"

F32 = 0
F64 = 1
UNEVEN_PERSISTENT = 1
UNEVEN_SORTED = 2

_lib = None
_lock = threading.Lock()

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_szp = ctypes.POINTER(ctypes.c_size_t)

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGNATURES = {
    "lxa_lu_factor": [_vp, _int, _i64, _i64, _vp, _i64, _vp, _vp],
    "lxa_lu_solve": [_vp, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int],
    "lxa_lu_small_solve": [_vp, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp],
    "lxa_chol_factor": [_vp, _int, _i64, _i64, _vp, _i64, _int, _vp],
    "lxa_chol_solve": [_vp, _int, _i64, _i64, _vp, _i64, _vp, _i64, _int],
    "lxa_chol_uneven_workspace_size": [_i64, _szp],
    "lxa_chol_uneven_solve": [
        _vp, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _int, _int, _dbl,
        _vp, _vp, _vp, _vp, ctypes.c_size_t,
    ],
    "lxa_chol_uneven_tune_solve": [
        _vp, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _int, _int, _dbl,
        _vp, _vp, _vp, _vp, ctypes.c_size_t, _int, _int,
    ],
    "lxa_chol_uneven_tuned_classes": [_vp, _vp],
    "lxa_chol_uneven_solve_noalloc": [
        _vp, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _int, _int, _dbl, _vp, _vp, _vp,
    ],
    "lxa_gemv_workspace_size": [_int, _i64, _i64, _i64, _int, _szp],
    "lxa_gemv": [_vp, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _int, _vp, ctypes.c_size_t],
    "lxa_solve_results": [_vp, _int, _i64, _i64, _vp, _i64, _vp, _i64, _vp],
    "lxa_stencil5_apply": [_vp, _int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64],
    "lxa_krylov_partials_size": [_i64, _i64, _int, _szp],
    "lxa_cg_begin": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_cg_alpha": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_cg_update": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _vp],
    "lxa_cg_residual": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _int, _vp, _vp, _vp],
    "lxa_cg_precond_gamma": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_cg_finish": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_cg_result": [_vp, _i64, _vp, _vp],
    "lxa_bicgstab_begin": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_bicgstab_direction": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _int, _vp],
    "lxa_bicgstab_alpha": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_bicgstab_s": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp],
    "lxa_bicgstab_omega": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_bicgstab_update": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_bicgstab_result": [_vp, _i64, _vp, _vp],
    "lxa_lsmr_begin": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_lsmr_init": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_lsmr_step_u": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_lsmr_step_v": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    "lxa_lsmr_step_x": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_lsmr_result": [_vp, _i64, _vp, _vp],
    "lxa_cg_stencil_workspace": [_i64, _i64, _i64, _szp],
    "lxa_gemv_uneven": [_vp, _i64, _vp, _int, _int, _int],
    "lxa_svd_workspace": [_int, _i64, _i64, _i64, _szp],
    "lxa_svd": [_vp, _int, _i64, _i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _int, _vp],
    "lxa_gmres_stencil_workspace": [_i64, _i64, _i64, _szp, _szp],
    "lxa_gmres_stencil_step": [_vp, _int, _i64, _i64, _i64, _vp, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _vp, _vp],
    "lxa_gmres_update_scaled": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_cg_stencil_direction": [_vp, _int, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _int, _vp, _vp, _vp],
    "lxa_cg_stencil_update": [_vp, _int, _int, _int, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _vp,
                              _vp],
    "lxa_gmres_init": [_vp, _i64, _vp, _vp],
    "lxa_gmres_cycle_begin": [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_gmres_arnoldi": [_vp, _int, _i64, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_gmres_update": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "lxa_gmres_residual": [_vp, _int, _int, _i64, _i64, _vp, _vp, _vp, _int, _vp, _vp, _vp],
    "lxa_gmres_cycle_end": [_vp, _i64, _i64, _int, _vp, _vp, _vp, _vp],
    "lxa_gmres_result": [_vp, _i64, _vp, _vp, _vp],
    "lxa_gemm_sub": [_vp, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64],
    "lxa_gemm_batched": [_vp, _int, _i64, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64],
    "lxa_lu_large_block": [],
    "lxa_lu_large_factor": [_vp, _int, _i64, _vp, _i64, _vp, _vp, _vp, _vp],
    "lxa_lu_large_perm": [_vp, _i64, _vp, _vp],
    "lxa_lu_large_permute": [_vp, _i64, _int, _vp, _i64, _vp, _i64, _vp, _int],
    "lxa_lu_large_trsv": [_vp, _int, _i64, _int, _vp, _i64, _vp, _vp, _i64, _vp],
    "lxa_lu_large_trsv_workspace": [_i64, _vp],
    "lxa_rbt_matrix": [_vp, _i64, _vp, _i64, _vp, _vp],
    "lxa_rbt_vectors": [_vp, _int, _i64, _int, _vp, _i64, _vp, _int],
    "lxa_gemv_f32a": [_vp, _i64, _i64, _int, _vp, _i64, _int, _vp, _i64, _vp, _i64, _int, _vp, _i64],
    "lxa_qr_block": [_int],
    "lxa_qr_work_elems": [_int, _i64, _i64, _i64p],
    "lxa_qr_factor": [_vp, _int, _i64, _vp, _int, _int, _int, _vp, _vp, _vp, _vp],
    "lxa_qr_apply": [_vp, _int, _i64, _vp, _int, _int, _vp, _vp, _int, _vp, _int],
    "lxa_qr_factor_step": [_vp, _int, _i64, _vp, _vp, _int, _int, _vp, _vp, _vp, _vp],
    "lxa_chol_blocked_factor": [_vp, _i64, _vp, _vp, _vp],
    "lxa_chol_blocked_solve": [_vp, _i64, _vp, _vp, _vp],
    "lxa_trsv": [_vp, _int, _i64, _vp, _int, _vp, _int, _vp, _int, _int, _int],
    "lxa_version": [],
    "lxa_last_error": [],
}
_RESTYPES = {"lxa_version": ctypes.c_char_p, "lxa_last_error": ctypes.c_char_p}


def declared_symbols() -> list[str]:
    """Names of every entry point this module binds."""
    return sorted(_SIGNATURES)


def load() -> ctypes.CDLL:
    """Loads the kernel library (once); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"lineax_amd kernel library not found at {LIB_PATH}; build it with "
                    "`python -c 'import __graft_entry__ as g; g.build()'`"
                )
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, argtypes in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = argtypes
                fn.restype = _RESTYPES.get(name, ctypes.c_int)
            _lib = lib
    return _lib


def version() -> str:
    return load().lxa_version().decode()


def _check(rc: int, name: str) -> None:
    if rc != 0:
        msg = load().lxa_last_error().decode()
        raise RuntimeError(f"{name} failed (status {rc}): {msg}")


def _dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return F32
    if dtype == torch.float64:
        return F64
    raise TypeError(f"lineax_amd kernels support float32 and float64, got {dtype}")


def _gpu(t: torch.Tensor, what: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor, got {type(t)}")
    if not t.is_cuda:
        raise ValueError(
            f"{what} is on {t.device}; lineax_amd solves run on the GPU (HIP) only"
        )


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _flat_batch(t: torch.Tensor, core: int) -> tuple[torch.Tensor, int, int]:
    """Views t as [batch, *core dims] with a uniform batch stride.

    Returns (tensor with contiguous core, batch, batch_stride_in_elements).
    A batch of identical (broadcast, stride-0) leading dims gives stride 0.
    """
    lead = t.shape[: t.dim() - core]
    batch = math.prod(lead)
    inner = t.shape[t.dim() - core:]
    inner_contig = True
    expect = 1
    for size, stride in zip(reversed(inner), reversed(t.stride()[t.dim() - core:])):
        if size != 1 and stride != expect:
            inner_contig = False
        expect *= size
    if inner_contig:
        lead_strides = t.stride()[: t.dim() - core]
        if all(s == 0 or sz == 1 for s, sz in zip(lead_strides, lead)):
            return t, batch, 0
        try:
            v = t.view(batch, *inner) if batch > 0 else t.reshape(0, *inner)
            return v, batch, (v.stride(0) if batch > 1 else math.prod(inner))
        except RuntimeError:
            pass
    v = t.reshape(batch, *inner).contiguous()
    return v, batch, math.prod(inner)


# ---------------------------------------------------------------------------
# LU
# ---------------------------------------------------------------------------


def lu_factor(matrix: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """getrf over leading batch dims: returns (lu, piv int32, info int32)."""
    _gpu(matrix, "matrix")
    *lead, n, n2 = matrix.shape
    if n != n2:
        raise ValueError(f"lu_factor needs square matrices, got {tuple(matrix.shape)}")
    dt = _dtype_code(matrix.dtype)
    lu = matrix.contiguous().clone()
    batch = math.prod(lead)
    piv = torch.empty((*lead, n), dtype=torch.int32, device=matrix.device)
    info = torch.empty(tuple(lead), dtype=torch.int32, device=matrix.device)
    if batch == 0 or n == 0:
        return lu, piv, info
    lib = load()
    _check(
        lib.lxa_lu_factor(_stream(lu), dt, batch, n, _ptr(lu), n * n, _ptr(piv), _ptr(info)),
        "lxa_lu_factor",
    )
    return lu, piv, info


def lu_solve(lu: torch.Tensor, piv: torch.Tensor, b: torch.Tensor, trans: int) -> torch.Tensor:
    """getrs: x with A x = b (trans 0) or A^T x = b (trans 1); batch dims broadcast."""
    _gpu(lu, "lu")
    _gpu(b, "vector")
    n = lu.shape[-1]
    if b.shape[-1] != n:
        raise ValueError(f"vector of size {b.shape[-1]} does not match operator of order {n}")
    if b.dtype != lu.dtype:
        raise TypeError(f"vector dtype {b.dtype} differs from factor dtype {lu.dtype}")
    dt = _dtype_code(lu.dtype)
    lead = torch.broadcast_shapes(lu.shape[:-2], piv.shape[:-1], b.shape[:-1])
    x = b.expand(*lead, n).contiguous().clone()
    batch = math.prod(lead)
    if batch == 0 or n == 0:
        return x
    luv, lb, lus = _flat_batch(lu.expand(*lead, n, n), 2)
    pvv, pb, pvs = _flat_batch(piv.to(torch.int32).expand(*lead, n), 1)
    lib = load()
    _check(
        lib.lxa_lu_solve(_stream(x), dt, batch, n, _ptr(luv), lus, _ptr(pvv), pvs, _ptr(x), n, int(trans)),
        "lxa_lu_solve",
    )
    return x


SMALL_LU_MAX = 64


def lu_small_solve(matrix: torch.Tensor, vector: torch.Tensor, want_factors: bool = False):
    """Fused getrf + getrs for orders <= 64 (lxa_lu_small_solve): returns
    (x, lu, piv, info); lu / piv are None unless ``want_factors``.  The matrix
    and vector batches must match (no broadcasting)."""
    _gpu(matrix, "matrix")
    _gpu(vector, "vector")
    *lead, n, n2 = matrix.shape
    if n != n2 or n > SMALL_LU_MAX:
        raise ValueError(f"lu_small_solve needs square matrices of order <= {SMALL_LU_MAX}, got {tuple(matrix.shape)}")
    if tuple(vector.shape) != (*lead, n):
        raise ValueError(f"vector batch {tuple(vector.shape)} does not match the matrix batch {tuple(matrix.shape)}")
    if vector.dtype != matrix.dtype:
        raise TypeError(f"vector dtype {vector.dtype} differs from matrix dtype {matrix.dtype}")
    dt = _dtype_code(matrix.dtype)
    a = matrix.contiguous()
    b = vector.contiguous()
    batch = math.prod(lead)
    dev = matrix.device
    x = torch.empty_like(b)
    lu = torch.empty_like(a) if want_factors else None
    piv = torch.empty((*lead, n), dtype=torch.int32, device=dev) if want_factors else None
    info = torch.empty(tuple(lead), dtype=torch.int32, device=dev)
    if batch == 0:
        return x, lu, piv, info
    _check(load().lxa_lu_small_solve(_stream(a), dt, batch, n, _ptr(a), n * n, _ptr(b), n, _ptr(x), n,
                                     _ptr(lu), _ptr(piv), _ptr(info)), "lxa_lu_small_solve")
    return x, lu, piv, info


# ---------------------------------------------------------------------------
# Cholesky
# ---------------------------------------------------------------------------


def chol_factor(matrix: torch.Tensor, negate: bool = False) -> tuple[torch.Tensor, torch.Tensor]:
    """potrf (upper) over leading batch dims: returns (U with zeros below, info)."""
    _gpu(matrix, "matrix")
    *lead, n, n2 = matrix.shape
    if n != n2:
        raise ValueError(f"chol_factor needs square matrices, got {tuple(matrix.shape)}")
    dt = _dtype_code(matrix.dtype)
    u = matrix.contiguous().clone()
    info = torch.empty(tuple(lead), dtype=torch.int32, device=matrix.device)
    batch = math.prod(lead)
    if batch == 0 or n == 0:
        return u, info
    lib = load()
    _check(
        lib.lxa_chol_factor(_stream(u), dt, batch, n, _ptr(u), n * n, int(bool(negate)), _ptr(info)),
        "lxa_chol_factor",
    )
    return u, info


def chol_solve(u: torch.Tensor, b: torch.Tensor, negate: bool = False) -> torch.Tensor:
    """potrs with U from chol_factor; batch dims broadcast."""
    _gpu(u, "factor")
    _gpu(b, "vector")
    n = u.shape[-1]
    if b.shape[-1] != n:
        raise ValueError(f"vector of size {b.shape[-1]} does not match operator of order {n}")
    if b.dtype != u.dtype:
        raise TypeError(f"vector dtype {b.dtype} differs from factor dtype {u.dtype}")
    dt = _dtype_code(u.dtype)
    lead = torch.broadcast_shapes(u.shape[:-2], b.shape[:-1])
    x = b.expand(*lead, n).contiguous().clone()
    batch = math.prod(lead)
    if batch == 0 or n == 0:
        return x
    uv, _, us = _flat_batch(u.expand(*lead, n, n), 2)
    lib = load()
    _check(
        lib.lxa_chol_solve(_stream(x), dt, batch, n, _ptr(uv), us, _ptr(x), n, int(bool(negate))),
        "lxa_chol_solve",
    )
    return x


def chol_uneven_solve(
    n: torch.Tensor,
    a_offset: torch.Tensor,
    a: torch.Tensor,
    b_offset: torch.Tensor,
    b: torch.Tensor,
    x: torch.Tensor,
    order: torch.Tensor,
    *,
    max_n: int,
    persistent: bool,
    max_refine: int,
    tol: float,
    info: torch.Tensor,
    iters: torch.Tensor | None = None,
    eta: torch.Tensor | None = None,
    workspace: torch.Tensor | None = None,
    noalloc: bool = False,
    sorted_order: bool = False,
    classes: tuple[int, int] | None = None,
) -> None:
    """Raw launch of lxa_chol_uneven_solve (see the header for semantics).

    Arguments are device tensors already laid out; ``order`` may list a
    subset of the systems (one size bucket, for instance).  ``noalloc``
    calls lxa_chol_uneven_solve_noalloc instead (queue counters in the
    library's own device array; no workspace).  ``sorted_order`` declares
    ``order`` sorted by n descending (LXA_UNEVEN_SORTED).  ``classes``
    = (max4, max8) runs lxa_chol_uneven_tune_solve with those tile-class
    boundaries (the autotuner's hook).
    """
    for t, what in ((n, "n"), (a_offset, "a_offset"), (a, "a"), (b_offset, "b_offset"), (b, "b"),
                    (x, "x"), (order, "order"), (info, "info")):
        _gpu(t, what)
    if n.dtype != torch.int32 or order.dtype != torch.int32 or info.dtype != torch.int32:
        raise TypeError("n, order and info must be int32")
    if a_offset.dtype != torch.int64 or b_offset.dtype != torch.int64:
        raise TypeError("offsets must be int64")
    if a.dtype != torch.float32 or b.dtype != torch.float32:
        raise TypeError("uneven batches hold float32 matrices and right-hand sides")
    xdt = _dtype_code(x.dtype)
    count = order.numel()
    lib = load()
    flags = (UNEVEN_PERSISTENT | (UNEVEN_SORTED if sorted_order else 0)) if persistent else 0
    if noalloc:
        _check(
            lib.lxa_chol_uneven_solve_noalloc(
                _stream(a), count, int(max_n), _ptr(n), _ptr(a_offset), _ptr(a), _ptr(b_offset), _ptr(b),
                _ptr(x), xdt, _ptr(order), flags, int(max_refine), float(tol), _ptr(info), _ptr(iters), _ptr(eta),
            ),
            "lxa_chol_uneven_solve_noalloc",
        )
        return
    ws_ptr, ws_bytes = None, 0
    if persistent or classes is not None:
        need = ctypes.c_size_t(0)
        _check(lib.lxa_chol_uneven_workspace_size(count, ctypes.byref(need)), "lxa_chol_uneven_workspace_size")
        if workspace is None or workspace.numel() * workspace.element_size() < need.value:
            workspace = torch.empty(int(need.value), dtype=torch.uint8, device=a.device)
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    if classes is not None:
        _check(
            lib.lxa_chol_uneven_tune_solve(
                _stream(a), count, int(max_n), _ptr(n), _ptr(a_offset), _ptr(a), _ptr(b_offset), _ptr(b),
                _ptr(x), xdt, _ptr(order), flags, int(max_refine), float(tol),
                _ptr(info), _ptr(iters), _ptr(eta), ws_ptr, ws_bytes, int(classes[0]), int(classes[1]),
            ),
            "lxa_chol_uneven_tune_solve",
        )
        return
    _check(
        lib.lxa_chol_uneven_solve(
            _stream(a), count, int(max_n), _ptr(n), _ptr(a_offset), _ptr(a), _ptr(b_offset), _ptr(b),
            _ptr(x), xdt, _ptr(order), flags, int(max_refine), float(tol),
            _ptr(info), _ptr(iters), _ptr(eta), ws_ptr, ws_bytes,
        ),
        "lxa_chol_uneven_solve",
    )


def uneven_tuned_classes() -> tuple[int, int]:
    """(max4, max8): the tile-class boundaries of the library's autotune table."""
    m4, m8 = ctypes.c_int(0), ctypes.c_int(0)
    _check(load().lxa_chol_uneven_tuned_classes(ctypes.byref(m4), ctypes.byref(m8)), "lxa_chol_uneven_tuned_classes")
    return int(m4.value), int(m8.value)


# ---------------------------------------------------------------------------
# Operators and results
# ---------------------------------------------------------------------------


def gemv(matrix: torch.Tensor, x: torch.Tensor, trans: bool = False) -> torch.Tensor:
    """y = A x or A^T x over broadcast leading batch dims."""
    _gpu(matrix, "matrix")
    _gpu(x, "vector")
    if x.dtype != matrix.dtype:
        raise TypeError(f"vector dtype {x.dtype} differs from matrix dtype {matrix.dtype}")
    dt = _dtype_code(matrix.dtype)
    m, n = matrix.shape[-2:]
    if m > 1 and n > 1 and matrix.stride(-2) == 1 and matrix.stride(-1) == m:
        # a transposed view (MatrixLinearOperator.transpose()): run the other
        # product on the contiguous storage instead of copying the matrix
        return gemv(matrix.mT, x, trans=not trans)
    kin, kout = (m, n) if trans else (n, m)
    if x.shape[-1] != kin:
        raise ValueError(f"vector of size {x.shape[-1]} does not match operator input size {kin}")
    lead = torch.broadcast_shapes(matrix.shape[:-2], x.shape[:-1])
    batch = math.prod(lead)
    y = torch.empty((*lead, kout), dtype=matrix.dtype, device=matrix.device)
    if batch == 0 or kout == 0:
        return y
    if kin == 0:
        return y.zero_()
    av, _, as_ = _flat_batch(matrix.expand(*lead, m, n), 2)
    xv, _, xs = _flat_batch(x.expand(*lead, kin), 1)
    lib = load()
    work, wbytes = None, 0
    if trans:
        need = ctypes.c_size_t(0)
        _check(lib.lxa_gemv_workspace_size(dt, batch, m, n, 1, ctypes.byref(need)), "lxa_gemv_workspace_size")
        wbytes = int(need.value)
        if wbytes:
            work = torch.empty(wbytes // 8, dtype=torch.float64, device=matrix.device)
    _check(
        lib.lxa_gemv(_stream(y), dt, batch, m, n, _ptr(av), as_, _ptr(xv), xs, _ptr(y), kout, int(bool(trans)),
                     _ptr(work) if work is not None else None, wbytes),
        "lxa_gemv",
    )
    return y


SVD_MAX_SWEEPS = 60


def svd(matrix: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Thin SVD of [*batch, m, n]: (u [*batch, m, k], s [*batch, k] descending,
    vt [*batch, k, n], sweeps [*batch] int32), k = min(m, n) (one-sided
    Jacobi, lxa_svd)."""
    _gpu(matrix, "matrix")
    dt = _dtype_code(matrix.dtype)
    m, n = matrix.shape[-2:]
    k = min(m, n)
    lead = tuple(matrix.shape[:-2])
    batch = math.prod(lead)
    dev = matrix.device
    u = torch.empty((*lead, m, k), dtype=matrix.dtype, device=dev)
    s = torch.empty((*lead, k), dtype=matrix.dtype, device=dev)
    vt = torch.empty((*lead, k, n), dtype=matrix.dtype, device=dev)
    sweeps = torch.zeros(lead, dtype=torch.int32, device=dev)
    if batch == 0 or k == 0:
        return u, s, vt, sweeps
    a = matrix.reshape(batch, m, n).contiguous()
    need = ctypes.c_size_t(0)
    _check(load().lxa_svd_workspace(dt, batch, m, n, ctypes.byref(need)), "lxa_svd_workspace")
    work = torch.empty(max(1, int(need.value) // matrix.element_size()), dtype=matrix.dtype, device=dev)
    _check(load().lxa_svd(_stream(a), dt, batch, m, n, _ptr(a), m * n, _ptr(u), _ptr(s), _ptr(vt), _ptr(work),
                          SVD_MAX_SWEEPS, _ptr(sweeps)), "lxa_svd")
    return u, s, vt, sweeps


def solve_results(solution: torch.Tensor, vector: torch.Tensor, result: torch.Tensor) -> torch.Tensor:
    """RESULTS post-processing (lineax/_solve.py:104-123) on [*batch, len] rows.

    ``result`` (int32, batch shape) is updated in place and returned.
    """
    _gpu(solution, "solution")
    _gpu(vector, "vector")
    _gpu(result, "result")
    if result.dtype != torch.int32 or not result.is_contiguous():
        raise TypeError("result must be a contiguous int32 tensor")
    if solution.dtype != vector.dtype:
        vector = vector.to(solution.dtype)
    dt = _dtype_code(solution.dtype)
    rows = result.numel()
    length = solution.numel() // rows if rows else 0
    sol = solution.reshape(rows, -1).contiguous()
    vec = vector.expand_as(solution).reshape(rows, -1).contiguous()
    lib = load()
    _check(
        lib.lxa_solve_results(_stream(result), dt, rows, length, _ptr(sol), sol.stride(0), _ptr(vec),
                              vec.stride(0), _ptr(result)),
        "lxa_solve_results",
    )
    return result


# ---------------------------------------------------------------------------
# Grid stencils
# ---------------------------------------------------------------------------


def stencil5(x: torch.Tensor, nx: int, ny: int, coeffs, out_dtype=None, out: torch.Tensor | None = None) -> torch.Tensor:
    """Five-point stencil on grid functions: x is [..., nx, ny] or [..., nx*ny].

    ``out`` (optional): a contiguous tensor of x's element count in the output
    dtype that receives the result (no allocation)."""
    _gpu(x, "vector")
    out_dtype = out.dtype if out is not None else (out_dtype or x.dtype)
    size = nx * ny
    if x.numel() % size:
        raise ValueError(f"vector of shape {tuple(x.shape)} is not a batch of {nx}x{ny} grids")
    rows = x.numel() // size
    xv = x.reshape(rows, size)
    # rows may be strided (e.g. one column of a Krylov basis); each row's grid
    # must be contiguous
    if xv.stride(1) != 1 or (rows > 1 and xv.stride(0) < size):
        xv = xv.contiguous()
    xstride = xv.stride(0) if rows > 1 else size
    if out is not None:
        _gpu(out, "out")
        if out.numel() != x.numel() or not out.is_contiguous() or out.device != x.device:
            raise ValueError("out must be a contiguous tensor of the vector's size on its device")
        x0, x1 = xv.data_ptr(), xv.data_ptr() + xv.numel() * xv.element_size()
        o0, o1 = out.data_ptr(), out.data_ptr() + out.numel() * out.element_size()
        if x0 < o1 and o0 < x1:
            raise ValueError("out must not overlap the vector (the stencil reads rows ahead of those it writes)")
        y = out.view(rows, size)
    else:
        y = torch.empty((rows, size), dtype=out_dtype, device=x.device)
    if rows == 0:
        return y.reshape(x.shape)
    c = (ctypes.c_double * 5)(*[float(v) for v in coeffs])
    _check(
        load().lxa_stencil5_apply(_stream(x), _dtype_code(x.dtype), _dtype_code(out_dtype), rows, nx, ny,
                                  ctypes.cast(c, ctypes.c_void_p), _ptr(xv), xstride, _ptr(y), size),
        "lxa_stencil5_apply",
    )
    return y.reshape(x.shape)


# ---------------------------------------------------------------------------
# Krylov building blocks
# ---------------------------------------------------------------------------


class KrylovParams(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("rcond", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("max_steps", ctypes.c_int64),
        ("max_steps_given", ctypes.c_int32),
        ("norm_kind", ctypes.c_int32),
        ("restart", ctypes.c_int32),
        ("stagnation_iters", ctypes.c_int32),
    ]


ROW_BYTES = 64  # lxa_krylov_row and lxa_gmres_row


class RowState:
    """Device array of per-row Krylov state (lxa_krylov_row / lxa_gmres_row)."""

    def __init__(self, rows: int, device):
        self.raw = torch.zeros((rows, ROW_BYTES), dtype=torch.uint8, device=device)
        self.f64 = self.raw.view(torch.float64)  # [rows, 8]
        self.i32 = self.raw.view(torch.int32)  # [rows, 16]

    @property
    def ptr(self):
        return self.raw.data_ptr()

    # lxa_krylov_row fields
    def step(self):
        return self.i32[:, 12]

    def active(self):
        return self.i32[:, 13]

    def result(self):
        return self.i32[:, 14]


class LsmrParams(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("conlim", ctypes.c_double),
        ("max_steps", ctypes.c_int64),
        ("max_steps_given", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


# lxa_lsmr_row: 27 doubles, then itn, istop, active, result as int32
LSMR_ROW_BYTES = 232
LSMR_F64 = {name: i for i, name in enumerate(
    "alpha beta zetabar alphabar rho rhobar cbar sbar betadd betad rhodold tautildeold thetatilde zeta delta "
    "normA2 maxrbar minrbar condA normr normAr normb normx normA c_hbar c_x c_h".split())}
LSMR_I32 = {"itn": 54, "istop": 55, "active": 56, "result": 57}


class LsmrState:
    """Device array of per-row LSMR state (lxa_lsmr_row)."""

    def __init__(self, rows: int, device):
        self.raw = torch.zeros((rows, LSMR_ROW_BYTES), dtype=torch.uint8, device=device)
        self.f64 = self.raw.view(torch.float64)  # [rows, 29]
        self.i32 = self.raw.view(torch.int32)  # [rows, 58]

    @property
    def ptr(self):
        return self.raw.data_ptr()

    def scalar(self, name):
        return self.f64[:, LSMR_F64[name]]

    def int_field(self, name):
        return self.i32[:, LSMR_I32[name]]


def partials_buffer(rows: int, n: int, restart: int, device) -> torch.Tensor:
    need = ctypes.c_size_t(0)
    _check(load().lxa_krylov_partials_size(rows, n, restart, ctypes.byref(need)), "lxa_krylov_partials_size")
    return torch.empty(max(8, int(need.value) // 8), dtype=torch.float64, device=device)


class StencilCGWorkspace:
    """Partials and coefficients of the fused stencil CG sweeps."""

    def __init__(self, rows: int, nx: int, ny: int, coeffs, device):
        pb = ctypes.c_size_t(0)
        _check(load().lxa_cg_stencil_workspace(rows, nx, ny, ctypes.byref(pb)), "lxa_cg_stencil_workspace")
        self.partials = torch.empty(max(1, int(pb.value) // 8), dtype=torch.float64, device=device)
        self._coeffs = (ctypes.c_double * 5)(*[float(c) for c in coeffs])
        self.coeffs = ctypes.cast(self._coeffs, ctypes.c_void_p).value


class StencilGMRESWorkspace:
    """Partials, step coefficients and coefficients of the fused stencil GMRES steps."""

    def __init__(self, rows: int, nx: int, ny: int, restart: int, coeffs, device):
        pb, cb = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _check(load().lxa_gmres_stencil_workspace(rows, nx, ny, ctypes.byref(pb), ctypes.byref(cb)),
               "lxa_gmres_stencil_workspace")
        self.partials = torch.empty(max(1, int(pb.value) // 8), dtype=torch.float64, device=device)
        self.coefs = torch.zeros(max(1, int(cb.value) // 8), dtype=torch.float64, device=device)
        self.colscale = torch.zeros((rows, restart + 1), dtype=torch.float64, device=device)
        self._coeffs = (ctypes.c_double * 5)(*[float(c) for c in coeffs])
        self.coeffs = ctypes.cast(self._coeffs, ctypes.c_void_p).value


def call(name: str, *args) -> None:
    """Invokes a Krylov entry point; tensors are passed as device pointers."""
    conv = []
    for a in args:
        if isinstance(a, torch.Tensor):
            conv.append(a.data_ptr())
        elif isinstance(a, (RowState, LsmrState)):
            conv.append(a.ptr)
        elif isinstance(a, (KrylovParams, LsmrParams)):
            conv.append(ctypes.cast(ctypes.pointer(a), ctypes.c_void_p))
        else:
            conv.append(a)
    _check(getattr(load(), name)(*conv), name)


def dtype_code(dtype: torch.dtype) -> int:
    return _dtype_code(dtype)


def current_stream(t: torch.Tensor) -> int:
    return _stream(t)


# ---------------------------------------------------------------------------
# Householder QR and triangular solves
# ---------------------------------------------------------------------------

# lxa_qr_desc (56 bytes): offsets in elements, then m, n, lda, ldb.
QR_DESC = np.dtype(
    [("a_offset", "<i8"), ("tau_offset", "<i8"), ("t_offset", "<i8"), ("w_offset", "<i8"), ("b_offset", "<i8"),
     ("m", "<i4"), ("n", "<i4"), ("lda", "<i4"), ("ldb", "<i4")]
)
assert QR_DESC.itemsize == 56


def qr_block(dtype: torch.dtype) -> int:
    """Panel width of the blocked QR (columns per compact-WY block)."""
    return int(load().lxa_qr_block(_dtype_code(dtype)))


def qr_work_elems(dtype: torch.dtype, m: int, n: int) -> int:
    out = ctypes.c_int64(0)
    _check(load().lxa_qr_work_elems(_dtype_code(dtype), int(m), int(n), ctypes.byref(out)), "lxa_qr_work_elems")
    return int(out.value)


# lxa_gemv_desc (48 bytes): device addresses of A (float32), x (float64),
# b (float32, or 0) and y (float64), then m, n, lda.
GEMV_DESC = np.dtype([("a", "<i8"), ("x", "<i8"), ("b", "<i8"), ("y", "<i8"), ("m", "<i4"), ("n", "<i4"),
                      ("lda", "<i4"), ("pad", "<i4")])
assert GEMV_DESC.itemsize == 48


def gemv_uneven_raw(descs: torch.Tensor, count: int, trans: bool, max_m: int, max_n: int, stream_of: torch.Tensor):
    """lxa_gemv_uneven on device descriptors (uint8 tensor of GEMV_DESC records)."""
    _gpu(descs, "descs")
    _check(load().lxa_gemv_uneven(_stream(stream_of), int(count), _ptr(descs), int(bool(trans)), int(max_m),
                                  int(max_n)), "lxa_gemv_uneven")


def descs_to_device(descs: np.ndarray, device) -> torch.Tensor:
    """Copies a QR_DESC array to device memory (uint8 tensor)."""
    if descs.dtype not in (QR_DESC, GEMV_DESC):
        raise TypeError("descriptor array must have dtype QR_DESC or GEMV_DESC")
    raw = torch.from_numpy(np.ascontiguousarray(descs).view(np.uint8).reshape(-1).copy())
    return host_to_device(raw, device)


def host_to_device(host: torch.Tensor, device) -> torch.Tensor:
    """Copies a small host tensor to the device without waiting for the
    current stream: the data goes through pinned memory (torch's caching host
    allocator keeps the pinned block until the copy has run)."""
    dev = torch.device(device)
    if dev.type != "cuda":
        return host.to(dev)
    return host.pin_memory().to(dev, non_blocking=True)


def qr_factor_raw(descs: torch.Tensor, count: int, max_m: int, max_n: int, max_k: int, a: torch.Tensor,
                  tau: torch.Tensor, tblk: torch.Tensor, work: torch.Tensor) -> None:
    """lxa_qr_factor on caller-laid-out storage (uneven batches)."""
    for t, what in ((descs, "descs"), (a, "a"), (tau, "tau"), (tblk, "tblk"), (work, "work")):
        _gpu(t, what)
    if not (a.dtype == tau.dtype == tblk.dtype == work.dtype):
        raise TypeError("a, tau, tblk and work must share one dtype")
    _check(
        load().lxa_qr_factor(_stream(a), _dtype_code(a.dtype), int(count), _ptr(descs), int(max_m), int(max_n),
                             int(max_k), _ptr(a), _ptr(tau), _ptr(tblk), _ptr(work)),
        "lxa_qr_factor",
    )


def qr_factor_step_raw(descs: torch.Tensor, count: int, kcol: torch.Tensor, max_rows: int, max_trail_cols: int,
                       a: torch.Tensor, tau: torch.Tensor, tblk: torch.Tensor, work: torch.Tensor) -> None:
    """One panel step of every system at its own column kcol[s] (int32, device)."""
    for t, what in ((descs, "descs"), (kcol, "kcol"), (a, "a"), (tau, "tau"), (tblk, "tblk"), (work, "work")):
        _gpu(t, what)
    if kcol.dtype != torch.int32:
        raise TypeError("kcol must be int32")
    _check(
        load().lxa_qr_factor_step(_stream(a), _dtype_code(a.dtype), int(count), _ptr(descs), _ptr(kcol),
                                  int(max_rows), int(max_trail_cols), _ptr(a), _ptr(tau), _ptr(tblk), _ptr(work)),
        "lxa_qr_factor_step",
    )


def qr_apply_raw(descs: torch.Tensor, count: int, max_m: int, max_k: int, a: torch.Tensor, tblk: torch.Tensor,
                 nrhs: int, b: torch.Tensor, trans: bool) -> None:
    for t, what in ((descs, "descs"), (a, "a"), (tblk, "tblk"), (b, "b")):
        _gpu(t, what)
    if not (a.dtype == tblk.dtype == b.dtype):
        raise TypeError("a, tblk and b must share one dtype")
    _check(
        load().lxa_qr_apply(_stream(b), _dtype_code(a.dtype), int(count), _ptr(descs), int(max_m), int(max_k),
                            _ptr(a), _ptr(tblk), int(nrhs), _ptr(b), int(bool(trans))),
        "lxa_qr_apply",
    )


def trsv_raw(descs: torch.Tensor, count: int, max_k: int, a: torch.Tensor, nrhs: int, b: torch.Tensor, *,
             lower: bool, trans: bool, unit_diag: bool) -> None:
    for t, what in ((descs, "descs"), (a, "a"), (b, "b")):
        _gpu(t, what)
    if a.dtype != b.dtype:
        raise TypeError("a and b must share one dtype")
    _check(
        load().lxa_trsv(_stream(b), _dtype_code(a.dtype), int(count), _ptr(descs), int(max_k), _ptr(a), int(nrhs),
                        _ptr(b), int(bool(lower)), int(bool(trans)), int(bool(unit_diag))),
        "lxa_trsv",
    )


def qr_factor(matrix: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """geqrf over leading batch dims of a [*batch, m, n] matrix.

    Returns (a, tau, tblk): a holds R on and above the diagonal and the
    Householder vectors below it (row-major, as geqrf's raw output
    transposed), tau is [*batch, min(m, n)], tblk is [*batch, panels, nb, nb]
    (the compact-WY T factors the apply kernel uses).
    """
    _gpu(matrix, "matrix")
    *lead, m, n = matrix.shape
    dt = matrix.dtype
    _dtype_code(dt)
    nb = qr_block(dt)
    k = min(m, n)
    panels = -(-k // nb)
    a = matrix.contiguous().clone()
    batch = math.prod(lead)
    tau = torch.zeros((*lead, k), dtype=dt, device=matrix.device)
    tblk = torch.zeros((*lead, panels, nb, nb), dtype=dt, device=matrix.device)
    if batch == 0 or k == 0:
        return a, tau, tblk
    welems = qr_work_elems(dt, m, n)
    work = torch.empty(max(1, batch * welems), dtype=dt, device=matrix.device)
    d = np.zeros(batch, dtype=QR_DESC)
    idx = np.arange(batch, dtype=np.int64)
    d["a_offset"] = idx * m * n
    d["tau_offset"] = idx * k
    d["t_offset"] = idx * panels * nb * nb
    d["w_offset"] = idx * welems
    d["m"], d["n"], d["lda"] = m, n, n
    qr_factor_raw(descs_to_device(d, matrix.device), batch, m, n, k, a, tau, tblk, work)
    return a, tau, tblk


def _solve_descs(lead, op_lead, m, n, nb, length, device):
    count = math.prod(lead)
    k = min(m, n)
    panels = -(-k // nb)
    op_count = math.prod(op_lead)
    opi = np.broadcast_to(np.arange(op_count, dtype=np.int64).reshape(op_lead), lead).reshape(-1)
    d = np.zeros(count, dtype=QR_DESC)
    d["a_offset"] = opi * m * n
    d["tau_offset"] = opi * k
    d["t_offset"] = opi * panels * nb * nb
    d["b_offset"] = np.arange(count, dtype=np.int64) * length
    d["m"], d["n"], d["lda"], d["ldb"] = m, n, n, length
    return descs_to_device(d, device), count


def qr_solve(a: torch.Tensor, tblk: torch.Tensor, b: torch.Tensor, transpose: bool) -> torch.Tensor:
    """Solves with the factors of qr_factor (a is [*batch, M, N], M >= N).

    transpose = False: least squares min |A x - b| (b of length M, x of
    length N): x = R^{-1} (Q^T b)[:N] (lineax/_solver/qr.py:87-92).
    transpose = True: minimum-norm solution of A^T x = b (b of length N, x of
    length M): x = Q [R^{-T} b; 0] (qr.py:79-86).
    """
    _gpu(a, "factor")
    _gpu(b, "vector")
    M, N = a.shape[-2:]
    if M < N:
        raise ValueError("qr_solve expects the factor of a tall (or square) matrix")
    if b.dtype != a.dtype:
        raise TypeError(f"vector dtype {b.dtype} differs from factor dtype {a.dtype}")
    length_in = N if transpose else M
    if b.shape[-1] != length_in:
        raise ValueError(f"vector of size {b.shape[-1]} does not match operator (expected {length_in})")
    op_lead = tuple(a.shape[:-2])
    lead = tuple(torch.broadcast_shapes(op_lead, b.shape[:-1]))
    nb = qr_block(a.dtype)
    descs, count = _solve_descs(lead, op_lead, M, N, nb, M, a.device)
    buf = torch.zeros((*lead, M), dtype=a.dtype, device=a.device)
    if count == 0:
        return buf[..., : (M if transpose else N)]
    if N == 0:
        return buf if transpose else buf[..., :0]
    if transpose:
        buf[..., :N] = b
        trsv_raw(descs, count, N, a.contiguous(), 1, buf, lower=False, trans=True, unit_diag=False)
        qr_apply_raw(descs, count, M, N, a.contiguous(), tblk.contiguous(), 1, buf, trans=False)
        return buf
    buf.copy_(b.expand(*lead, M))
    qr_apply_raw(descs, count, M, N, a.contiguous(), tblk.contiguous(), 1, buf, trans=True)
    trsv_raw(descs, count, N, a.contiguous(), 1, buf, lower=False, trans=False, unit_diag=False)
    return buf[..., :N].contiguous()


def tri_solve(matrix: torch.Tensor, b: torch.Tensor, *, lower: bool, trans: bool, unit_diag: bool) -> torch.Tensor:
    """solve_triangular over broadcast batch dims (matrix [*batch, n, n])."""
    _gpu(matrix, "matrix")
    _gpu(b, "vector")
    n = matrix.shape[-1]
    if matrix.shape[-2] != n:
        raise ValueError("triangular solves need square matrices")
    if b.shape[-1] != n:
        raise ValueError(f"vector of size {b.shape[-1]} does not match operator of order {n}")
    if b.dtype != matrix.dtype:
        raise TypeError(f"vector dtype {b.dtype} differs from matrix dtype {matrix.dtype}")
    op_lead = tuple(matrix.shape[:-2])
    lead = tuple(torch.broadcast_shapes(op_lead, b.shape[:-1]))
    x = b.expand(*lead, n).contiguous().clone()
    if math.prod(lead) == 0 or n == 0:
        return x
    descs, count = _solve_descs(lead, op_lead, n, n, qr_block(matrix.dtype), n, matrix.device)
    trsv_raw(descs, count, n, matrix.contiguous(), 1, x, lower=lower, trans=trans, unit_diag=unit_diag)
    return x


# ---------------------------------------------------------------------------
# Matrix-core update and the large-LU building blocks
# ---------------------------------------------------------------------------

GEMM_F32 = 0
GEMM_BF16 = 1
GEMM_F16 = 2
GEMM_BF16X3 = 3
GEMM_TRANS_A, GEMM_TRANS_B = 256, 512  # LXA_GEMM_TRANS_A / _B: operand stored transposed
GEMM_MODES = {"fp32": GEMM_F32, "bf16": GEMM_BF16, "fp16": GEMM_F16, "bf16x3": GEMM_BF16X3}


def _f32_2d(t: torch.Tensor, what: str) -> None:
    _gpu(t, what)
    if t.dtype != torch.float32 or t.dim() != 2 or t.stride(1) != 1:
        raise TypeError(f"{what} must be a float32 matrix with unit column stride")


def gemm_sub(c: torch.Tensor, a: torch.Tensor, b: torch.Tensor, mode: str = "fp32") -> torch.Tensor:
    """C <- C - A B in place on the matrix cores (float32 storage)."""
    for t, what in ((a, "a"), (b, "b"), (c, "c")):
        _f32_2d(t, what)
    m, k = a.shape
    k2, n = b.shape
    if k2 != k or tuple(c.shape) != (m, n):
        raise ValueError(f"shape mismatch: {tuple(a.shape)} x {tuple(b.shape)} -> {tuple(c.shape)}")
    _check(load().lxa_gemm_sub(_stream(c), GEMM_MODES[mode], m, n, k, _ptr(a), a.stride(0), _ptr(b), b.stride(0),
                               _ptr(c), c.stride(0)), "lxa_gemm_sub")
    return c


GEMM_MAX_BATCH = 65535


def gemm(a: torch.Tensor, b: torch.Tensor, mode: str = "fp32") -> torch.Tensor:
    """A @ B of float32 matrices or batches of them (leading dimensions
    broadcast as in torch.matmul) on the matrix cores, one launch per 65535
    systems: the systems of the batch are the grid's third dimension."""
    for t, what in ((a, "a"), (b, "b")):
        _gpu(t, what)
        if t.dtype != torch.float32 or t.dim() < 2:
            raise TypeError(f"{what} must be a float32 matrix or batch of matrices")
    m, k = a.shape[-2:]
    k2, n = b.shape[-2:]
    if k2 != k:
        raise ValueError(f"shape mismatch: {tuple(a.shape)} @ {tuple(b.shape)}")
    batch_shape = torch.broadcast_shapes(a.shape[:-2], b.shape[:-2])
    count = int(np.prod(batch_shape)) if batch_shape else 1
    c = torch.empty(*batch_shape, m, n, dtype=torch.float32, device=a.device)
    if c.numel() == 0:
        return c
    if k == 0:
        return c.zero_()

    def flat(t):
        # (count, rows, cols) with one unit stride among the last two (a
        # transposed layout is read in place in float32 mode); a broadcast
        # operand keeps batch stride 0.  Returns (view, batch stride, leading
        # dimension, transposed).
        if t.shape[:-2] == batch_shape or not batch_shape:
            t = t.reshape(-1, *t.shape[-2:]) if t.dim() > 2 else t.unsqueeze(0)
        elif t.dim() == 2 or all(d == 1 for d in t.shape[:-2]):
            t = t.reshape(1, *t.shape[-2:])
        else:
            t = t.expand(*batch_shape, *t.shape[-2:]).contiguous().reshape(-1, *t.shape[-2:])
        rows, cols = t.shape[-2:]
        if t.stride(-1) == 1 and t.stride(-2) >= cols:
            trans, ld = False, t.stride(-2)
        elif mode == "fp32" and t.stride(-2) == 1 and t.stride(-1) >= rows:
            trans, ld = True, t.stride(-1)
        else:
            t = t.contiguous()
            trans, ld = False, cols
        return t, (t.stride(0) if t.shape[0] > 1 else 0), ld, trans

    af, sa, lda, ta = flat(a)
    bf, sb, ldb, tb = flat(b)
    cf = c.reshape(-1, m, n)
    code = GEMM_MODES[mode] | (GEMM_TRANS_A if ta else 0) | (GEMM_TRANS_B if tb else 0)
    lib = load()
    for s0 in range(0, count, GEMM_MAX_BATCH):
        cnt = min(GEMM_MAX_BATCH, count - s0)
        _check(lib.lxa_gemm_batched(_stream(c), code, cnt, m, n, k, _ptr(af[s0 if sa else 0]), lda, sa,
                                    _ptr(bf[s0 if sb else 0]), ldb, sb, _ptr(cf[s0]), n, m * n), "lxa_gemm_batched")
    return c


def lu_large_block() -> int:
    return int(load().lxa_lu_large_block())


def lu_large_factor(a: torch.Tensor, mode: str, linv: torch.Tensor, uinv: torch.Tensor, piv: torch.Tensor,
                    info: torch.Tensor) -> None:
    _f32_2d(a, "a")
    n = a.shape[0]
    if a.shape[1] != n:
        raise ValueError("lu_large_factor needs a square matrix")
    if piv.dtype != torch.int32 or piv.numel() < n:
        raise TypeError("piv must be int32 with n entries")
    _check(load().lxa_lu_large_factor(_stream(a), GEMM_MODES[mode], n, _ptr(a), a.stride(0), _ptr(linv), _ptr(uinv),
                                      _ptr(piv), _ptr(info)), "lxa_lu_large_factor")


def lu_large_perm(piv: torch.Tensor) -> torch.Tensor:
    """The factorisation's row permutation (perm[i] = original row at position i)."""
    _gpu(piv, "piv")
    perm = torch.empty_like(piv)
    _check(load().lxa_lu_large_perm(_stream(piv), piv.shape[0], _ptr(piv), _ptr(perm)), "lxa_lu_large_perm")
    return perm


def lu_large_permute(x: torch.Tensor, perm: torch.Tensor, inverse: bool) -> torch.Tensor:
    """P x (inverse=False) or P^T x of the rows [nrhs, n] of x, into a new tensor."""
    _gpu(x, "x")
    _gpu(perm, "perm")
    y = torch.empty_like(x)
    _check(load().lxa_lu_large_permute(_stream(x), x.shape[1], x.shape[0], _ptr(x), x.stride(0), _ptr(y), y.stride(0),
                                       _ptr(perm), int(bool(inverse))), "lxa_lu_large_permute")
    return y


LU_TRSV_FAILED_WORD = 1  # the workspace word a timed-out block wait sets


def lu_large_trsv_workspace(n: int, device) -> torch.Tensor:
    """Zeroed int32 workspace of lxa_lu_large_trsv for order n (16-byte aligned)."""
    words = ctypes.c_int64(0)
    _check(load().lxa_lu_large_trsv_workspace(int(n), ctypes.byref(words)), "lxa_lu_large_trsv_workspace")
    return torch.zeros(int(words.value), dtype=torch.int32, device=device)


def lu_large_trsv(kind: int, lu: torch.Tensor, inv: torch.Tensor, x: torch.Tensor, sync: torch.Tensor) -> None:
    """Block triangular solve in place on x [nrhs <= 8, ldx] (float32)."""
    _f32_2d(lu, "lu")
    _f32_2d(x, "x")
    n = lu.shape[0]
    _check(load().lxa_lu_large_trsv(_stream(x), int(kind), n, x.shape[0], _ptr(lu), lu.stride(0), _ptr(inv), _ptr(x),
                                    x.stride(0), _ptr(sync)), "lxa_lu_large_trsv")


def rbt_matrix(a: torch.Tensor, du: torch.Tensor, dv: torch.Tensor) -> None:
    _f32_2d(a, "a")
    _check(load().lxa_rbt_matrix(_stream(a), a.shape[0], _ptr(a), a.stride(0), _ptr(du), _ptr(dv)), "lxa_rbt_matrix")


def rbt_vectors(x: torch.Tensor, d: torch.Tensor, form: int) -> None:
    _gpu(x, "x")
    if x.dim() != 2 or x.stride(1) != 1:
        raise TypeError("x must be [nrhs, n] with unit column stride")
    _check(load().lxa_rbt_vectors(_stream(x), _dtype_code(x.dtype), x.shape[1], x.shape[0], _ptr(x), x.stride(0),
                                  _ptr(d), int(form)), "lxa_rbt_vectors")


def gemv_f32a(a: torch.Tensor, x: torch.Tensor, out_dtype: torch.dtype, b: torch.Tensor | None = None) -> torch.Tensor:
    """A x (or b - A x) for float32 A [m, k] and x [nrhs, k]; FP64 accumulation."""
    _f32_2d(a, "a")
    _gpu(x, "x")
    m, k = a.shape
    x = x if (x.dim() == 2 and x.stride(1) == 1) else x.reshape(-1, k).contiguous()
    if x.shape[1] != k:
        raise ValueError(f"x has {x.shape[1]} columns, expected {k}")
    nrhs = x.shape[0]
    y = torch.empty((nrhs, m), dtype=out_dtype, device=a.device)
    if b is not None:
        b = b.to(out_dtype).reshape(nrhs, m).contiguous()
    _check(load().lxa_gemv_f32a(_stream(y), m, k, nrhs, _ptr(a), a.stride(0), _dtype_code(x.dtype), _ptr(x),
                                x.stride(0), _ptr(b), (b.stride(0) if b is not None else m), _dtype_code(out_dtype),
                                _ptr(y), y.stride(0)), "lxa_gemv_f32a")
    return y


def chol_blocked_factor_raw(descs: torch.Tensor, count: int, a: torch.Tensor, info: torch.Tensor) -> None:
    """Blocked Cholesky of an uneven batch (orders <= 1024), float32, in place."""
    for t, what in ((descs, "descs"), (a, "a"), (info, "info")):
        _gpu(t, what)
    if a.dtype != torch.float32 or info.dtype != torch.int32:
        raise TypeError("a must be float32 and info int32")
    _check(load().lxa_chol_blocked_factor(_stream(a), int(count), _ptr(descs), _ptr(a), _ptr(info)),
           "lxa_chol_blocked_factor")


def chol_blocked_solve_raw(descs: torch.Tensor, count: int, u: torch.Tensor, b: torch.Tensor) -> None:
    for t, what in ((descs, "descs"), (u, "u"), (b, "b")):
        _gpu(t, what)
    if u.dtype != torch.float32 or b.dtype != torch.float32:
        raise TypeError("u and b must be float32")
    _check(load().lxa_chol_blocked_solve(_stream(b), int(count), _ptr(descs), _ptr(u), _ptr(b)),
           "lxa_chol_blocked_solve")
