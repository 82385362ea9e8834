"""ORACLE — test infrastructure only.

A CPU restatement of lineax's solver hot path in float64, used to check the
GPU kernels.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU
baseline leg may import it; the product path (``lineax_amd``) never does.

Dense factorisations call the C restatement of the LAPACK routines that
lineax reaches through JAX (``oracle/lxo_dense.c``, built by
``oracle/Makefile``).  The Krylov solvers are restated here line by line
from the reference:

  - ``linear_solve_results``: lineax/_solve.py:104-123
  - ``lu_factor``/``lu_solve``: lineax/_solver/lu.py:43-66
  - ``cho_factor``/``cho_solve``: lineax/_solver/cholesky.py:43-78
  - ``qr_factor``/``qr_solve``: lineax/_solver/qr.py:55-94
  - ``cg``: lineax/_solver/cg.py:114-227
  - ``gmres``: lineax/_solver/gmres.py:106-413
  - ``resolve_rcond``: lineax/_misc.py:30-38
  - ``max_norm``/``two_norm``: lineax/_norm.py:59-139
"""

from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "liblxo.so"

SUCCESSFUL, MAX_STEPS_REACHED, SINGULAR, BREAKDOWN, STAGNATION, CONLIM, NONFINITE_INPUT = range(7)

_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.lxo_getrf.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip]
        L.lxo_getrf.restype = ctypes.c_int
        L.lxo_getrs.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip, dp, ctypes.c_int]
        L.lxo_potrf_upper.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.lxo_potrf_upper.restype = ctypes.c_int
        L.lxo_potrs_upper.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp]
        L.lxo_geqrf.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp]
        L.lxo_ormqr_vec.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp, ctypes.c_int]
        L.lxo_trsv_upper.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int]
        L.lxo_chol_uneven.argtypes = [
            ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
        ]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64c(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# ---------------------------------------------------------------------------
# Dense factorisations (single systems; batches loop in Python)
# ---------------------------------------------------------------------------


def lu_factor(a):
    """getrf: (lu, piv, info) with 0-based interchanges (jax.scipy convention)."""
    lu = _f64c(a).copy()
    n = lu.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib().lxo_getrf(n, _dp(lu), n, _ip(piv))
    return lu, piv, info


def lu_solve(lu, piv, b, trans=0):
    x = _f64c(b).copy()
    lib().lxo_getrs(lu.shape[0], _dp(_f64c(lu)), lu.shape[0], _ip(np.ascontiguousarray(piv, np.int32)), _dp(x), trans)
    return x


def cho_factor(a, negate=False):
    """potrf 'U' of A (or -A): (U with zeros below, info); NaN factor on failure."""
    u = _f64c(a).copy()
    if negate:
        u = -u
    n = u.shape[0]
    info = lib().lxo_potrf_upper(n, _dp(u), n)
    if info:
        return np.full_like(u, np.nan), info
    return np.triu(u), 0


def cho_solve(u, b, negate=False):
    x = _f64c(b).copy()
    lib().lxo_potrs_upper(u.shape[0], _dp(_f64c(u)), u.shape[0], _dp(x))
    return -x if negate else x


def qr_factor(a):
    """lineax QR.init (qr.py:55-65): geqrf of A, or of A^T when A is wide.

    Returns (a_raw [max(m,n), min(m,n)] with R above and reflectors below,
    taus, transposed).
    """
    a = _f64c(a)
    m, n = a.shape
    transpose = n > m
    if transpose:
        a = np.ascontiguousarray(a.T)
        m, n = n, m
    h = a.copy()
    tau = np.zeros(min(m, n))
    lib().lxo_geqrf(m, n, _dp(h), n, _dp(tau))
    return h, tau, transpose


def qr_solve(h, tau, transpose, b):
    """lineax QR.compute (qr.py:67-94).

    Not transposed: least squares x = R^{-1} (Q^T b)[:n].
    Transposed (operator was wide): minimum norm x = Q (R^{-T} b ; 0).
    """
    n_full, n_min = h.shape
    h = np.ascontiguousarray(h)
    if transpose:
        y = _f64c(b).copy()
        lib().lxo_trsv_upper(n_min, _dp(h), n_min, _dp(y), 1)
        c = np.concatenate([y, np.zeros(n_full - n_min)])
        lib().lxo_ormqr_vec(n_full, n_min, _dp(h), n_min, _dp(np.ascontiguousarray(tau)), _dp(c), 0)
        return c
    c = _f64c(b).copy()
    lib().lxo_ormqr_vec(n_full, n_min, _dp(h), n_min, _dp(np.ascontiguousarray(tau)), _dp(c), 1)
    x = np.ascontiguousarray(c[:n_min])
    lib().lxo_trsv_upper(n_min, _dp(h), n_min, _dp(x), 0)
    return x


def chol_uneven(ns, a_off, a32, b_off, b32, threads=1):
    """float64 Cholesky solves of an uneven batch held as float32 (see lxo_dense.c)."""
    ns = np.ascontiguousarray(ns, np.int32)
    a_off = np.ascontiguousarray(a_off, np.int64)
    b_off = np.ascontiguousarray(b_off, np.int64)
    a32 = np.ascontiguousarray(a32, np.float32)
    b32 = np.ascontiguousarray(b32, np.float32)
    x = np.zeros(b32.shape[0], np.float64)
    info = np.zeros(ns.shape[0], np.int32)
    lib().lxo_chol_uneven(
        ns.shape[0], ns.ctypes.data, a_off.ctypes.data, a32.ctypes.data, b_off.ctypes.data, b32.ctypes.data,
        x.ctypes.data, info.ctypes.data, int(threads),
    )
    return x, info


# ---------------------------------------------------------------------------
# linear_solve semantics
# ---------------------------------------------------------------------------


def normal_solve(a, b):
    """lx.Normal(lx.Cholesky()) (lineax/_solver/normal.py:105-140): for m >= n
    x = (A^T A)^-1 A^T b, else x = A^T (A A^T)^-1 b, each Gram matrix factored
    by potrf 'U' and solved by potrs; float64."""
    a = _f64c(a)
    b = np.asarray(b, np.float64)
    m, n = a.shape
    if m >= n:
        u, _ = cho_factor(a.T @ a)
        return cho_solve(u, a.T @ b)
    u, _ = cho_factor(a @ a.T)
    return a.T @ cho_solve(u, b)


def svd_solve(a, b, rcond=None):
    """lx.SVD() (lineax/_solver/svd.py:49-78): thin SVD, singular values kept
    when s > rcond * s_max (rcond default 2 eps max(m, n),
    lineax/_misc.py:30-38), x = V diag(1/s) U^T b; returns (x, rank)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m, n = a.shape
    u, s, vt = np.linalg.svd(a, full_matrices=False)
    rc = 2 * np.finfo(np.float64).eps * max(m, n) if rcond is None else rcond
    cutoff = rc * s[0] if s.size else 0.0
    mask = s > cutoff
    s_inv = np.where(mask, 1.0 / np.where(mask, s, 1.0), 0.0)
    return vt.T @ (s_inv * (u.T @ b)), int(mask.sum())


def linear_solve_results(solution, vector, result):
    """lineax/_solve.py:104-123 for one system."""
    if result == SUCCESSFUL and not np.all(np.isfinite(solution)):
        result = SINGULAR
    if result == SINGULAR and not np.all(np.isfinite(vector)):
        result = NONFINITE_INPUT
    return result


def solve_dense(kind, a, b, *, negate=False):
    """linear_solve(MatrixLinearOperator(a), b, LU()/Cholesky()/QR()) -> (x, result)."""
    if kind == "lu":
        lu, piv, _ = lu_factor(a)
        x = lu_solve(lu, piv, b)
    elif kind == "cholesky":
        u, _ = cho_factor(a, negate=negate)
        x = cho_solve(u, b, negate=negate)
    elif kind == "qr":
        h, tau, t = qr_factor(a)
        x = qr_solve(h, tau, t, b)
    else:
        raise ValueError(kind)
    return x, linear_solve_results(x, b, SUCCESSFUL)


def normwise_backward_error(a, x, b):
    """eta = |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) (Rigal-Gaches), in float64."""
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64)
    b = np.asarray(b, np.float64)
    r = b - a @ x
    denom = np.abs(a).sum(axis=-1).max() * np.abs(x).max() + np.abs(b).max()
    return float(np.abs(r).max() / denom) if denom > 0 else float(np.abs(r).max())


def lstsq_backward_error(a, x, b):
    """eta_LS = |A^T (b - A x)|_inf / (|A|_1 (|A|_inf |x|_inf + |b|_inf)), in
    float64: the normal-equations residual of a least-squares solution
    (zero at the exact minimiser, about m u_f64 at its FP64 rounding), scaled
    so that eta_LS <= 1 for any x; as |A^T r|_inf <= |A|_1 |r|_inf, it never
    exceeds the normwise backward error of A x = b.
    The accuracy measure for the least-squares half of the request loop, in
    the spirit of LAPACK's xQRT17 ratio |r^T A| / (|A| |B| max(m, n) eps)."""
    a = np.asarray(a, np.float64)
    x = np.asarray(x, np.float64)
    b = np.asarray(b, np.float64)
    g = a.T @ (b - a @ x)
    denom = np.abs(a).sum(axis=0).max() * (np.abs(a).sum(axis=-1).max() * np.abs(x).max() + np.abs(b).max())
    return float(np.abs(g).max() / denom) if denom > 0 else float(np.abs(g).max())


def relative_error(x, x_ref):
    x = np.asarray(x, np.float64)
    x_ref = np.asarray(x_ref, np.float64)
    d = np.abs(x_ref).max()
    return float(np.abs(x - x_ref).max() / d) if d > 0 else float(np.abs(x - x_ref).max())


# ---------------------------------------------------------------------------
# Norms and rcond
# ---------------------------------------------------------------------------


def max_norm(x):
    x = np.asarray(x)
    return np.max(np.abs(x)) if x.size else 0.0


def two_norm(x):
    x = np.asarray(x)
    if x.size == 1:
        return float(np.abs(x.reshape(())))
    return float(np.sqrt(np.dot(x.ravel(), x.ravel())))


def resolve_rcond(n, m, dtype):
    return 2 * np.finfo(dtype).eps * max(n, m)


# ---------------------------------------------------------------------------
# CG (lineax/_solver/cg.py:114-227)
# ---------------------------------------------------------------------------


def cg(mv, b, rtol, atol, *, max_steps=None, stabilise_every=10, y0=None, precond=None, norm=max_norm,
       dtype=np.float64, negate=False, return_trace=False):
    """Returns (y, result, num_steps).  ``mv`` is the operator's matvec.

    ``negate``: the operator is NSD; CG runs on -A and negates the solution
    (cg.py:100-101, 224-225).
    """
    b = np.asarray(b, dtype)
    op = (lambda v: -mv(v)) if negate else mv
    M = precond if precond is not None else (lambda v: v)
    size = b.size
    steps_cap = 10 * size if max_steps is None else max_steps
    y = np.zeros_like(b) if y0 is None else np.asarray(y0, dtype).copy()
    r = b - op(y)
    p = M(r)
    gamma = np.dot(p, r)
    rcond = resolve_rcond(size, size, dtype)
    diff = np.full_like(y, np.inf)
    has_scale = not (atol == 0 and rtol == 0)
    b_scale = atol + rtol * np.abs(b)
    step = 0
    trace = []

    def not_converged(r, diff, y):
        if not has_scale:
            return True
        y_scale = atol + rtol * np.abs(y)
        with np.errstate(divide="ignore", invalid="ignore"):
            n1 = norm(r / b_scale)
            n2 = norm(diff / y_scale)
        return bool(n1 > 1) or bool(n2 > 1)

    while bool(gamma > 0) and step < steps_cap and not_converged(r, diff, y):
        q = op(p)
        ip = np.dot(q, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            alpha = gamma / ip
        if not bool(abs(ip) > 100 * rcond * abs(gamma)):
            alpha = np.nan
        diff = alpha * p
        y = y + diff
        step += 1
        if stabilise_every == 1:
            r = b - op(y)
        elif stabilise_every is None:
            r = r - alpha * q
        elif step % stabilise_every == 0:
            r = b - op(y)
        else:
            r = r - alpha * q
        z = M(r)
        gamma_prev = gamma
        gamma = np.dot(z, r)
        with np.errstate(divide="ignore", invalid="ignore"):
            beta = gamma / gamma_prev
        p = z + beta * p
        if return_trace:
            trace.append(float(np.sqrt(np.dot(r, r))))

    if max_steps is None:
        result = SINGULAR if step == steps_cap else SUCCESSFUL
    elif has_scale:
        result = MAX_STEPS_REACHED if step == steps_cap else SUCCESSFUL
    else:
        result = SUCCESSFUL
    if negate:
        y = -y
    result = linear_solve_results(y, b, result)
    if return_trace:
        return y, result, step, trace
    return y, result, step


# ---------------------------------------------------------------------------
# GMRES (lineax/_solver/gmres.py:106-413)
# ---------------------------------------------------------------------------


# ---------------------------------------------------------------------------
# BiCGStab (lineax/_solver/bicgstab.py:78-205)
# ---------------------------------------------------------------------------


def bicgstab(mv, b, rtol, atol, *, max_steps=None, y0=None, precond=None, norm=max_norm, dtype=np.float64,
             x64=True):
    """Returns (y, result, num_steps), following bicgstab.py line by line:
    r0 = b - A y0 (103); breakdown test (113-119: exact zeros under
    jax_enable_x64, else < 1e-16); cond_fun (121-140); body (143-167) with
    right preconditioning; initial carry alpha = omega = rho = 1, diff = inf
    (169-180); RESULTS (185-197)."""
    b = np.asarray(b, dtype)
    M = precond if precond is not None else (lambda v: v)
    size = b.size
    steps_cap = 10 * size if max_steps is None else max_steps
    has_scale = not (atol == 0 and rtol == 0)
    b_scale = atol + rtol * np.abs(b)
    y = np.zeros_like(b) if y0 is None else np.asarray(y0, dtype).copy()
    r0 = b - mv(y)

    def breakdown(omega, alpha, rho):
        if x64:
            return bool(omega == 0.0) or bool(alpha == 0.0) or bool(rho == 0.0)
        return bool(omega < 1e-16) or bool(alpha < 1e-16) or bool(rho < 1e-16)

    def not_converged(r, diff, y):
        if not has_scale:
            return True
        y_scale = atol + rtol * np.abs(y)
        with np.errstate(divide="ignore", invalid="ignore"):
            n1 = norm(r / b_scale)
            n2 = norm(diff / y_scale)
        return bool(n1 > 1) or bool(n2 > 1)

    r = r0.copy()
    alpha = omega = rho = 1.0
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    diff = np.full_like(y, np.inf)
    step = 0
    while (not breakdown(omega, alpha, rho)) and not_converged(r, diff, y) and step < steps_cap:
        with np.errstate(divide="ignore", invalid="ignore"):
            rho_new = np.dot(r0, r)
            beta = (rho_new / rho) * (alpha / omega)
            p = r + beta * (p - omega * v)
            x = M(p)
            v = mv(x)
            alpha = rho_new / np.dot(r0, v)
            s = r - alpha * v
            z = M(s)
            t = mv(z)
            omega = np.dot(s, t) / np.dot(t, t)
            diff = alpha * x + omega * z
            y = y + diff
            r = s - omega * t
        rho = rho_new
        step += 1

    if max_steps is None:
        result = SINGULAR if step == steps_cap else SUCCESSFUL
    elif has_scale:
        result = MAX_STEPS_REACHED if step == steps_cap else SUCCESSFUL
    else:
        result = SUCCESSFUL
    if breakdown(omega, alpha, rho) and not_converged(r, diff, y):
        result = BREAKDOWN
    return y, result, step


# ---------------------------------------------------------------------------
# LSMR (lineax/_solver/lsmr.py:94-409)
# ---------------------------------------------------------------------------


def _givens(a, b):
    """lsmr.py:361-409: c, s, r with [c -s; s c] [a; b] = [r; 0]."""
    if a == 0.0 or b == 0.0:
        if b == 0.0:
            return float(np.sign(a)), 0.0, abs(a)
        return 0.0, float(np.sign(b)), abs(b)
    if abs(b) > abs(a):
        tau = a / b
        s = np.sign(b) / np.sqrt(1.0 + tau * tau)
        c = s * tau
        return c, s, b / s
    tau = b / a
    c = np.sign(a) / np.sqrt(1.0 + tau * tau)
    s = c * tau
    return c, s, a / c


def lsmr(mv, rmv, b, n, rtol, atol, *, max_steps=None, conlim=1e8, x0=None, stats=False):
    """Returns (x, result, num_steps, istop), following lsmr.py line by line
    (damp = 0, two_norm): initialisation 92-187, the loop body 189-333,
    RESULTS 347-362.  ``mv``/``rmv`` apply A and A^T.  With ``stats`` a
    fifth element holds the estimates of lsmr.py:334-342 (norm_r, norm_Ar,
    norm_A = sqrt(normA2), cond_A)."""
    b = np.asarray(b, np.float64)
    m = b.size
    nrm = lambda z: float(np.sqrt(np.dot(z, z)))
    has_scale = not (atol == 0 and rtol == 0)
    steps_cap = min(m, n) * 10 if max_steps is None else max_steps
    x = np.zeros(n) if x0 is None else np.asarray(x0, np.float64).copy()
    u = b - mv(x)
    normb = nrm(b)
    beta = nrm(u)
    if beta != 0.0:
        u = u / beta
        v = rmv(u)
        alpha = nrm(v)
    else:
        v = np.zeros(n)
        alpha = 0.0
    v = v / (1.0 if alpha == 0.0 else alpha)
    h = v.copy()
    hbar = np.zeros(n)
    itn = 0
    zetabar, alphabar, rho, rhobar, cbar, sbar = alpha * beta, alpha, 1.0, 1.0, 1.0, 0.0
    betadd, betad, rhodold, tautildeold, thetatilde, zeta, delta = beta, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0
    normA2, maxrbar, minrbar, condA = alpha ** 2, 0.0, np.finfo(np.float64).max, 1.0
    istop, normr, normAr = 0, beta, alpha * beta
    if alpha == 0.0:
        istop = 2
    if beta == 0.0:
        istop = 1
    while istop == 0:
        itn += 1
        u = -alpha * u + mv(v)
        beta = nrm(u)
        if beta != 0.0:
            u = u / beta
            v = -beta * v + rmv(u)
            alpha = nrm(v)
            v = v / (1.0 if alpha == 0.0 else alpha)
        chat, shat, alphahat = _givens(alphabar, 0.0)
        rhoold = rho
        c, s, rho = _givens(alphahat, beta)
        thetanew = s * alpha
        alphabar = c * alpha
        rhobarold = rhobar
        zetaold = zeta
        thetabar = sbar * rho
        rhotemp = cbar * rho
        cbar, sbar, rhobar = _givens(cbar * rho, thetanew)
        zeta = cbar * zetabar
        zetabar = -sbar * zetabar
        hbar = hbar * -(thetabar * rho / (rhoold * rhobarold)) + h
        x = x + (zeta / (rho * rhobar)) * hbar
        h = h * -(thetanew / rho) + v
        betaacute = chat * betadd
        betacheck = -shat * betadd
        betahat = c * betaacute
        betadd = -s * betaacute
        thetatildeold = thetatilde
        ctildeold, stildeold, rhotildeold = _givens(rhodold, thetabar)
        thetatilde = stildeold * rhobar
        rhodold = ctildeold * rhobar
        betad = -stildeold * betad + ctildeold * betahat
        tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold
        taud = (zeta - thetatilde * tautildeold) / rhodold
        delta = delta + betacheck ** 2
        normr = np.sqrt(delta + (betad - taud) ** 2 + betadd ** 2)
        normA2 = normA2 + beta ** 2
        normA = np.sqrt(normA2)
        normA2 = normA2 + alpha ** 2
        maxrbar = max(maxrbar, rhobarold)
        if itn > 1:
            minrbar = min(minrbar, rhobarold)
        condA = max(maxrbar, rhotemp) / min(minrbar, rhotemp)
        normAr = abs(zetabar)
        normx = nrm(x)
        if itn >= steps_cap:
            istop = 4
        if condA > conlim:
            istop = 3
        if normAr < atol + rtol * (normA * normr):
            istop = 2
        if normr < atol + rtol * (normA * normx + normb):
            istop = 1
    if max_steps is None:
        result = SINGULAR if itn == steps_cap else SUCCESSFUL
    elif has_scale:
        result = MAX_STEPS_REACHED if itn == steps_cap else SUCCESSFUL
    else:
        result = SUCCESSFUL
    if istop < 3:
        result = SUCCESSFUL
    if istop == 3:
        result = CONLIM
    if stats:
        return x, result, itn, istop, {"norm_r": normr, "norm_Ar": normAr, "norm_A": float(np.sqrt(normA2)),
                                       "cond_A": condA}
    return x, result, itn, istop


def _normalise(x, eps=None):
    nrm = two_norm(x)
    if eps is None:
        eps = np.finfo(np.asarray(x).dtype).eps
    breakdown = nrm < eps
    safe = np.inf if breakdown else nrm
    return x / safe, nrm, breakdown


def _hessenberg_lstsq(coeff_mat, r_norm):
    """z = linear_solve(MatrixLinearOperator(coeff_mat.T), beta e1, QR(), throw=False)."""
    h = coeff_mat.T  # (restart + 1, restart)
    beta = np.zeros(h.shape[0])
    beta[0] = r_norm
    raw, tau, transposed = qr_factor(h)
    return qr_solve(raw, tau, transposed, beta)


def gmres(mv, b, rtol, atol, *, max_steps=None, restart=20, stagnation_iters=20, y0=None, precond=None,
          norm=max_norm, dtype=np.float64):
    """Returns (y, result, num_steps) with lineax's restart/breakdown/stagnation rules."""
    b = np.asarray(b, dtype)
    M = precond if precond is not None else (lambda v: v)
    has_scale = not (atol == 0 and rtol == 0)
    b_scale = atol + rtol * np.abs(b)
    size = b.size
    steps_cap = 10 * size if max_steps is None else max_steps
    restart = min(restart, size)
    eps = np.finfo(dtype).eps

    def not_converged(r, diff, y):
        if not has_scale:
            return True
        y_scale = atol + rtol * np.abs(y)
        with np.errstate(divide="ignore", invalid="ignore"):
            n1 = norm(r / b_scale)
            n2 = norm(diff / y_scale)
        return bool(n1 > 1) or bool(n2 > 1)

    def cycle(y, r, first_pass):
        if first_pass:
            y_new, diff, breakdown = y, np.full_like(y, np.inf), False
        else:
            r_normalised, r_norm, breakdown = _normalise(r)
            basis = np.zeros((size, restart + 1), dtype)
            basis[:, 0] = r_normalised
            coeff = np.eye(restart, restart + 1, dtype=dtype)
            j = 0
            while j < restart and not breakdown:
                w = M(mv(basis[:, j]))
                step_norm = two_norm(w)
                proj = basis.T @ w
                w_new = w - basis @ proj
                w_normed, w_norm, breakdown = _normalise(w_new, eps=step_norm * eps)
                basis[:, j + 1] = w_normed
                proj[j + 1] = w_norm
                coeff[j, :] = proj
                j += 1
            z = _hessenberg_lstsq(coeff, r_norm)
            diff = basis[:, :-1] @ z
            y_new = y + diff
        r_new = M(b - mv(y_new))
        return y_new, r_new, breakdown, diff

    y = np.zeros_like(b) if y0 is None else np.asarray(y0, dtype).copy()
    r = np.zeros_like(b)
    breakdown = False
    deferred = False
    diff = np.full_like(y, np.inf)
    r_min = np.inf
    step = 0
    stag = 0

    def cond():
        out = (not deferred) and stag < stagnation_iters
        out = out and not_converged(r, diff, y)
        out = out and step < steps_cap
        return out or step == 0

    while cond():
        y_new, r_new, breakdown_new, diff_new = cycle(y, r, step == 0)
        r_new_norm = norm(r_new)
        if (r_new_norm - r_min) < 0:
            stag = 0
        else:
            stag += 1
        r_min = min(r_new_norm, r_min)
        deferred = breakdown
        breakdown = breakdown_new
        y, r, diff = y_new, r_new, diff_new
        step += 1

    if max_steps is None:
        result = SINGULAR if step == steps_cap else SUCCESSFUL
    elif has_scale:
        result = MAX_STEPS_REACHED if step == steps_cap else SUCCESSFUL
    else:
        result = SUCCESSFUL
    if stag >= stagnation_iters:
        result = STAGNATION
    if deferred and not_converged(r, diff, y):
        result = BREAKDOWN
    result = linear_solve_results(y, b, result)
    return y, result, step


# ---------------------------------------------------------------------------
# Grid operators of the Krylov workloads (fixture definitions)
# ---------------------------------------------------------------------------


def stencil_apply(x, coeffs):
    """Five-point stencil on a (nx, ny) grid, zero Dirichlet boundary;
    coeffs = (center, west j-1, east j+1, south i-1, north i+1)."""
    c, w, e, s, n = coeffs
    x = np.asarray(x, np.float64)
    xp = np.pad(x, ((1, 1), (1, 1)))
    return (c * x + w * xp[1:-1, :-2] + e * xp[1:-1, 2:] + s * xp[:-2, 1:-1] + n * xp[2:, 1:-1])


def stencil_matrix(nx, ny, coeffs):
    """The same operator as a scipy.sparse CSR matrix (row-major grid order)."""
    import scipy.sparse as sp

    c, w, e, s, n = coeffs
    N = nx * ny
    idx = np.arange(N).reshape(nx, ny)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(N, c)]
    for di, dj, v in ((0, -1, w), (0, 1, e), (-1, 0, s), (1, 0, n)):
        src = idx[max(0, -di): nx - max(0, di), max(0, -dj): ny - max(0, dj)]
        dst = idx[max(0, di): nx + min(0, di), max(0, dj): ny + min(0, dj)]
        rows.append(src.ravel())
        cols.append(dst.ravel())
        vals.append(np.full(src.size, v))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
