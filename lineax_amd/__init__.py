"""lineax_amd: lineax's linear-solve API on hand-written gfx950 (MI355X) kernels.

The public names mirror lineax (lineax/__init__.py).  Arrays are torch
tensors on a HIP device; solves run in the kernels of
``lineax_amd/_lib/liblineax_amd.so`` (C ABI: ``include/lineax_amd.h``).
"""

from ._misc import resolve_rcond
from ._norm import max_norm, rms_norm, sum_squares, tree_dot, two_norm
from ._operator import (
    AbstractLinearOperator,
    AddLinearOperator,
    ComposedLinearOperator,
    DivLinearOperator,
    FunctionLinearOperator,
    IdentityLinearOperator,
    JacobianLinearOperator,
    MatrixLinearOperator,
    MulLinearOperator,
    NegLinearOperator,
    TaggedLinearOperator,
    conj,
    diagonal,
    has_unit_diagonal,
    is_diagonal,
    is_lower_triangular,
    is_negative_semidefinite,
    is_positive_semidefinite,
    is_symmetric,
    is_tridiagonal,
    is_upper_triangular,
    linearise,
    materialise,
)
from ._solution import RESULTS, Solution
from ._solve import AbstractLinearSolver, AutoLinearSolver, invert, linear_solve, sentinel
from ._solver import (
    CG,
    GMRES,
    LSMR,
    LU,
    QR,
    SVD,
    BiCGStab,
    Cholesky,
    Diagonal,
    Normal,
    Triangular,
    Tridiagonal,
)
from ._structure import TensorStruct
from ._tags import (
    diagonal_tag,
    lower_triangular_tag,
    negative_semidefinite_tag,
    positive_semidefinite_tag,
    symmetric_tag,
    transpose_tags,
    tridiagonal_tag,
    unit_diagonal_tag,
    upper_triangular_tag,
)

__version__ = "0.1.0"
