"""The expected tangents of the derivative tests, pinned on the CPU.

tests.helpers.dx_closed_form (the derivative of the least-squares and
minimum-norm solutions, which lineax's JVP rule lineax/_solve.py:151-268
computes) against central differences of numpy's own solutions, float64.
"""

import numpy as np
import pytest

from tests.helpers import dx_closed_form, lstsq


@pytest.mark.parametrize("shape", [(9, 9), (14, 6), (5, 11), (30, 12), (8, 21)])
def test_closed_form_matches_central_differences(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.standard_normal(shape) + (3 * np.eye(*shape) if shape[0] == shape[1] else 0)
    b, da, db = rng.standard_normal(shape[0]), rng.standard_normal(shape), rng.standard_normal(shape[0])
    h = 1e-6
    fd = (lstsq(a + h * da, b + h * db) - lstsq(a - h * da, b - h * db)) / (2 * h)
    np.testing.assert_allclose(dx_closed_form(a, b, da, db), fd, rtol=1e-6, atol=1e-7)


def test_closed_form_catches_a_dropped_term():
    # the tall formula without its A'^T r term misses the finite difference
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((14, 6)), rng.standard_normal(14)
    da, db = rng.standard_normal((14, 6)), rng.standard_normal(14)
    x = lstsq(a, b)
    wrong = np.linalg.solve(a.T @ a, a.T @ (db - da @ x))
    h = 1e-6
    fd = (lstsq(a + h * da, b + h * db) - lstsq(a - h * da, b - h * db)) / (2 * h)
    assert np.max(np.abs(wrong - fd)) > 1e-3 * np.max(np.abs(fd))
