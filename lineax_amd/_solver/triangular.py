"""Triangular solver (lineax/_solver/triangular.py) — kernels pending."""

from __future__ import annotations

from .unsupported import _Unsupported


class Triangular(_Unsupported):
    _reference = "lineax/_solver/triangular.py"

    def assume_full_rank(self):
        return True
