from .bicgstab import BiCGStab
from .cg import CG
from .cholesky import Cholesky
from .diagonal import Diagonal
from .gmres import GMRES
from .lsmr import LSMR
from .lu import LU
from .normal import Normal
from .qr import QR
from .svd import SVD
from .triangular import Triangular
from .unsupported import Tridiagonal

__all__ = [
    "BiCGStab", "CG", "Cholesky", "Diagonal", "GMRES", "LSMR", "LU", "Normal", "QR", "SVD", "Triangular",
    "Tridiagonal",
]
