"""Shared test helpers: matrix construction in the style of the reference tests.

``well_conditioned`` follows lineax tests/helpers.py:30-83 (random normal
matrices, resampled until well conditioned).  For definite matrices of any
order it uses M M^T / n + 0.1 I (condition number below ~50) instead of
resampling M M^T, whose condition number grows like n^2.
"""

from __future__ import annotations

import numpy as np


def well_conditioned(rng, n, *, kind="general", cond_max=None):
    if kind in ("psd", "nsd"):
        m = rng.standard_normal((n, n))
        a = m @ m.T / n + 0.1 * np.eye(n)
        return a if kind == "psd" else -a
    cond_max = cond_max or max(1000.0, 10.0 * n * n)
    while True:
        m = rng.standard_normal((n, n))
        if np.linalg.cond(m) < cond_max:
            return m


def poisson_1d(n):
    """lineax tests/helpers.py:101-107: -2 on the diagonal, 1 off it (NSD)."""
    return -2 * np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)


def lstsq(a, b):
    """numpy's least-squares (tall) / minimum-norm (wide) solution, float64."""
    return np.linalg.lstsq(a, b, rcond=None)[0]


def dx_closed_form(a, b, da, db):
    """Tangent of x = A^+ b for a full-rank A (square, tall or wide): the
    derivative of the least-squares / minimum-norm solution."""
    m, n = a.shape
    x = lstsq(a, b)
    if m == n:
        return np.linalg.solve(a, db - da @ x)
    if m > n:
        r = b - a @ x
        return np.linalg.solve(a.T @ a, da.T @ r + a.T @ (db - da @ x))
    pinv = np.linalg.pinv(a)
    return pinv @ (db - da @ x) + (np.eye(n) - pinv @ a) @ (da.T @ np.linalg.solve(a @ a.T, b))
