"""Small numerical helpers (lineax/_misc.py)."""

from __future__ import annotations

import torch


def finfo_eps(dtype: torch.dtype) -> float:
    return torch.finfo(dtype).eps


def resolve_rcond(rcond, n: int, m: int, dtype: torch.dtype) -> float:
    """Default relative cutoff 2 * eps * max(n, m) (lineax/_misc.py:30-38).

    CG's breakdown guard uses it: the step is discarded (NaN) unless
    |<Ap, p>| > 100 * rcond * |gamma|.
    """
    if rcond is None:
        return 2 * finfo_eps(dtype) * max(n, m)
    return finfo_eps(dtype) if rcond < 0 else float(rcond)
