"""CG solver (lineax/_solver/cg.py) — kernels pending."""

from __future__ import annotations

from .unsupported import _Unsupported


class CG(_Unsupported):
    _reference = "lineax/_solver/cg.py"

    def assume_full_rank(self):
        return True
