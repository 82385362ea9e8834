"""Host-side behaviour that needs no GPU: API surface, validation, dispatch,
tag rules, and the C ABI library's exported symbols."""

import ctypes
import pathlib
import re

import numpy as np
import pytest
import torch

import lineax_amd as lx
from lineax_amd import _kernels as K

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _header_symbols():
    text = (ROOT / "include" / "lineax_amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(lxa_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    if not K.LIB_PATH.exists():
        pytest.skip("kernel library not built")
    lib = ctypes.CDLL(str(K.LIB_PATH))
    declared = _header_symbols()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"{name} missing from {K.LIB_PATH.name}"
    # Every declared entry point is bound by the Python boundary, and vice versa.
    assert set(declared) == set(K.declared_symbols())
    assert K.version().startswith("lineax_amd")


def test_results_codes_and_messages():
    assert [lx.RESULTS.successful, lx.RESULTS.max_steps_reached, lx.RESULTS.singular, lx.RESULTS.breakdown,
            lx.RESULTS.stagnation, lx.RESULTS.conlim, lx.RESULTS.nonfinite_input] == list(range(7))
    assert lx.RESULTS[0] == ""
    assert "non-finite" in lx.RESULTS[lx.RESULTS.singular]
    assert "non-finite" in lx.RESULTS[torch.tensor([0, 6, 0])]


def test_transpose_tags():
    t = lx.transpose_tags(frozenset([lx.lower_triangular_tag, lx.unit_diagonal_tag]))
    assert t == frozenset([lx.upper_triangular_tag, lx.unit_diagonal_tag])
    sym = frozenset([lx.symmetric_tag, lx.lower_triangular_tag])
    assert lx.transpose_tags(sym) == sym


def test_predicates_and_auto_dispatch():
    m = torch.eye(3)
    plain = lx.MatrixLinearOperator(m)
    psd = lx.MatrixLinearOperator(m, lx.positive_semidefinite_tag)
    nsd = lx.MatrixLinearOperator(-m, lx.negative_semidefinite_tag)
    wide = lx.MatrixLinearOperator(torch.ones(2, 3))
    assert lx.is_symmetric(psd) and lx.is_positive_semidefinite(psd)
    assert lx.is_negative_semidefinite(-psd) and lx.is_positive_semidefinite(-nsd)
    auto = lx.AutoLinearSolver(well_posed=True)
    assert isinstance(auto.select_solver(plain), lx.LU)
    assert isinstance(auto.select_solver(psd), lx.Cholesky)
    assert isinstance(auto.select_solver(nsd), lx.Cholesky)
    assert isinstance(lx.AutoLinearSolver(well_posed=None).select_solver(wide), lx.QR)
    with pytest.raises(ValueError, match="non-square"):
        auto.select_solver(wide)
    tri = lx.MatrixLinearOperator(m, lx.lower_triangular_tag)
    assert isinstance(auto.select_solver(tri), lx.Triangular)
    assert isinstance(lx.AutoLinearSolver(well_posed=False).select_solver(plain), lx.SVD)
    assert lx.AutoLinearSolver(well_posed=False).assume_full_rank() is False


def test_psd_tag_needs_square_structure():
    with pytest.raises(ValueError, match="matching input and output"):
        lx.MatrixLinearOperator(torch.ones(2, 3), lx.positive_semidefinite_tag)


def test_linear_solve_validation():
    with pytest.raises(ValueError, match="raw tensor"):
        lx.linear_solve(torch.eye(2), torch.ones(2))
    op = lx.MatrixLinearOperator(torch.eye(3))
    with pytest.raises(ValueError, match="structures do not match"):
        lx.linear_solve(op, torch.ones(4))
    with pytest.raises(ValueError, match="structures do not match"):
        lx.linear_solve(op, torch.ones(3, dtype=torch.float64))


def test_cpu_tensors_fail_loudly():
    # No CPU fallback: kernels are the only compute path.
    op = lx.MatrixLinearOperator(torch.eye(3) * 2)
    with pytest.raises((ValueError, RuntimeError)):
        lx.linear_solve(op, torch.ones(3), lx.LU())
    with pytest.raises((ValueError, RuntimeError)):
        lx.linear_solve(lx.MatrixLinearOperator(torch.eye(3), lx.positive_semidefinite_tag), torch.ones(3),
                        lx.Cholesky())


def test_solver_structure_errors():
    with pytest.raises(ValueError, match="square"):
        lx.LU().init(lx.MatrixLinearOperator(torch.ones(2, 3)), {})
    with pytest.raises(ValueError, match="positive or negative definite"):
        lx.Cholesky().init(lx.MatrixLinearOperator(torch.eye(2)), {})


def test_identity_operator_short_circuit():
    ident = lx.IdentityLinearOperator(lx.TensorStruct((3,), torch.float32))
    v = torch.arange(3.0)
    sol = lx.linear_solve(ident, v)
    assert sol.value is v and int(sol.result) == 0


def test_unsupported_solvers_raise():
    with pytest.raises(NotImplementedError, match="tridiagonal.py"):
        lx.Tridiagonal().init(lx.MatrixLinearOperator(torch.eye(2), lx.tridiagonal_tag), {})


def test_solvers_refuse_host_tensors():
    # no CPU fallback: a host operator is an error, not a silent CPU solve
    with pytest.raises(ValueError, match="GPU"):
        lx.SVD().init(lx.MatrixLinearOperator(torch.eye(2)), {})


def test_norms():
    x = torch.tensor([3.0, -4.0])
    assert float(lx.two_norm(x)) == 5.0
    assert float(lx.max_norm(x)) == 4.0
    assert float(lx.tree_dot(x, x)) == 25.0
    assert float(lx.two_norm(torch.tensor([-2.0]))) == 2.0


def test_rcond_default():
    assert lx.resolve_rcond(None, 10, 4, torch.float32) == 2 * np.finfo(np.float32).eps * 10


def test_workloads_are_seeded():
    from lineax_amd import workloads as W

    a1, b1 = W.small_lu_batch(batch=4, n=8, seed=0)
    a2, b2 = W.small_lu_batch(batch=4, n=8, seed=0)
    np.testing.assert_array_equal(a1, a2)
    assert a1.dtype == np.float32 and a1.shape == (4, 8, 8)
    u1 = W.uneven_spd_batch(count=20, n_max=24, seed=1)
    u2 = W.uneven_spd_batch(count=20, n_max=24, seed=1)
    assert u1.digest() == u2.digest()
    for i in range(u1.count):
        a = u1.matrix(i).astype(np.float64)
        np.testing.assert_array_equal(a, a.T)
        assert np.linalg.eigvalsh(a).min() > 0.09


def test_uneven_batch_rejects_bad_layout():
    from lineax_amd.batch import UnevenBatch

    with pytest.raises(ValueError, match="orders"):
        UnevenBatch([300], [0], np.zeros(90000, np.float32), [0], np.zeros(300, np.float32), device="cpu")
    with pytest.raises(ValueError, match="run past"):
        UnevenBatch([4], [10], np.zeros(16, np.float32), [0], np.zeros(4, np.float32), device="cpu")


def test_uneven_workspace_minimum_refused_on_host():
    """A persistent uneven solve with a workspace below the reported size is
    refused with LXA_EINVAL before anything reaches the device."""
    if not K.LIB_PATH.exists():
        pytest.skip("kernel library not built")
    lib = K.load()
    need = ctypes.c_size_t(0)
    assert lib.lxa_chol_uneven_workspace_size(10, ctypes.byref(need)) == 0
    assert need.value >= 12  # three 4-byte queue counters
    fake = ctypes.c_void_p(4096)  # never dereferenced: validation fails first
    rc = lib.lxa_chol_uneven_solve(None, 10, 256, fake, fake, fake, fake, fake, fake, K.F64, fake,
                                   K.UNEVEN_PERSISTENT, 2, 1e-12, fake, None, None, fake, 8)
    assert rc == 1  # LXA_EINVAL
    assert "workspace" in lib.lxa_last_error().decode()


def test_large_lu_trsv_timeout_flag_raises():
    """A block triangular solve that gave up waiting (device flag set) makes
    the large-LU solve raise instead of returning a wrong answer."""
    from lineax_amd.large_lu import LargeLU

    fac = LargeLU.__new__(LargeLU)
    fac.trsv_failed = torch.zeros(1, dtype=torch.int32)
    fac.check_trsv()
    fac.trsv_failed = torch.ones(1, dtype=torch.int32)
    with pytest.raises(RuntimeError, match="spin limit"):
        fac.check_trsv()
