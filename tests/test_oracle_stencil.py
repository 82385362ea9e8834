"""The stencil fixtures: matrix-free apply equals the sparse matrix, the
Poisson operator is SPD with the known eigenpairs, and the upwind
convection-diffusion coefficients follow the documented formula."""

import numpy as np

from lineax_amd.stencil import convection_diffusion_coeffs
from oracle import lineax_oracle as O


def test_apply_matches_matrix(rng):
    coeffs = (4.1, -1.2, -0.9, -1.3, -0.7)
    x = rng.standard_normal((5, 7))
    A = O.stencil_matrix(5, 7, coeffs)
    np.testing.assert_allclose(O.stencil_apply(x, coeffs).ravel(), A @ x.ravel(), rtol=1e-14)


def test_poisson_eigenpair():
    n = 9
    coeffs = (4.0, -1.0, -1.0, -1.0, -1.0)
    k, l_ = 2, 3
    i = np.arange(1, n + 1)
    v = np.outer(np.sin(np.pi * k * i / (n + 1)), np.sin(np.pi * l_ * i / (n + 1)))
    lam = 4 - 2 * np.cos(np.pi * k / (n + 1)) - 2 * np.cos(np.pi * l_ / (n + 1))
    np.testing.assert_allclose(O.stencil_apply(v, coeffs), lam * v, atol=1e-13)
    A = O.stencil_matrix(n, n, coeffs).toarray()
    assert np.allclose(A, A.T) and np.linalg.eigvalsh(A).min() > 0


def test_convection_diffusion_coefficients():
    n = 1023
    c, w, e, s, nn = convection_diffusion_coeffs(n, peclet=10.0, velocity=(1.0, 1.0))
    h = 1.0 / (n + 1)
    assert abs(c - (0.4 + 2 * h)) < 1e-15
    assert abs(w - (-0.1 - h)) < 1e-15 and abs(s - (-0.1 - h)) < 1e-15
    assert e == -0.1 and nn == -0.1
    # Row sums: the interior rows of an upwind discretisation sum to zero.
    assert abs(c + w + e + s + nn) < 1e-15
