"""QR solver (lineax/_solver/qr.py) — kernels pending."""

from __future__ import annotations

from .unsupported import _Unsupported


class QR(_Unsupported):
    _reference = "lineax/_solver/qr.py"

    def assume_full_rank(self):
        return True
