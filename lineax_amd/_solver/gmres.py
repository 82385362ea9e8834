"""GMRES solver (lineax/_solver/gmres.py) — kernels pending."""

from __future__ import annotations

from .unsupported import _Unsupported


class GMRES(_Unsupported):
    _reference = "lineax/_solver/gmres.py"

    def assume_full_rank(self):
        return True
