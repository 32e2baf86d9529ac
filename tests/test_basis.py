"""LGL basis, split derivative matrix and the CK2N54 coefficients, pinned by
closed forms, quadrature/derivative exactness, the SBP property and the order
conditions -- for the oracle and for the product library alike."""
import numpy as np
import pytest

import trixiatmo_hip as T


def product_basis(n):
    x, w, D, Ds = T.lgl_basis(n - 1)
    return x, w, D, Ds


@pytest.mark.parametrize("src", ["oracle", "product"])
def test_lgl_closed_forms(oracle, src):
    get = (lambda n: oracle.lgl(n)) if src == "oracle" else (lambda n: product_basis(n)[:2])
    x, w = get(4)
    np.testing.assert_allclose(x, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], rtol=0, atol=2e-16)
    np.testing.assert_allclose(w, [1 / 6, 5 / 6, 5 / 6, 1 / 6], rtol=0, atol=2e-16)
    x, w = get(5)
    np.testing.assert_allclose(x, [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], rtol=0, atol=2e-16)
    np.testing.assert_allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=0, atol=4e-16)


@pytest.mark.parametrize("src", ["oracle", "product"])
@pytest.mark.parametrize("n", range(2, 10))
def test_quadrature_and_derivative_exactness(oracle, src, n):
    if src == "oracle":
        x, w = oracle.lgl(n)
        D, Ds = oracle.dmat(n), oracle.dsplit(n)
    else:
        x, w, D, Ds = product_basis(n)
    # LGL quadrature is exact up to degree 2n - 3
    for deg in range(2 * n - 2):
        exact = (1 - (-1) ** (deg + 1)) / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-14
    # D is exact on polynomials up to degree n - 1
    for deg in range(n):
        d = deg * x ** (deg - 1) if deg > 0 else np.zeros(n)
        np.testing.assert_allclose(D @ x ** deg, d, atol=5e-13 * max(1, n * n))
    # SBP: M D + D^T M = diag(-1, 0, ..., 0, 1)
    M = np.diag(w)
    Bm = np.zeros((n, n))
    Bm[0, 0], Bm[-1, -1] = -1, 1
    np.testing.assert_allclose(M @ D + D.T @ M, Bm, atol=5e-14 * n)
    # split operator has a vanishing diagonal and equals 2D + diag(1/w1, 0.., -1/wn)
    np.testing.assert_allclose(np.diag(Ds), 0, atol=5e-13 * n)
    ref = 2 * D
    ref[0, 0] += 1 / w[0]
    ref[-1, -1] -= 1 / w[-1]
    np.testing.assert_allclose(Ds - np.diag(np.diag(Ds)), ref - np.diag(np.diag(ref)), atol=1e-13)


@pytest.mark.parametrize("n", range(2, 9))
def test_product_basis_equals_oracle(oracle, n):
    x, w, D, Ds = product_basis(n)
    xo, wo = oracle.lgl(n)
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, wo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(D, oracle.dmat(n), rtol=0, atol=1e-12)


def test_ck2n54_order_conditions(oracle):
    """The 2N scheme applied to y' = z y gives y1 = R(z) y0 with a degree-5
    polynomial R; fourth order means R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24 + O(z^5),
    and the stage times c_s must be the consistent abscissae."""
    A, B, c = oracle.ck2n54()

    def R(z):
        y, k = 1.0 + 0j, 0j
        for s in range(5):
            k = A[s] * k + z * y
            y = y + B[s] * k
        return y

    zs = np.array([0.3 * np.exp(2j * np.pi * q / 6) for q in range(6)])
    V = np.vander(zs, 6, increasing=True)
    coef = np.linalg.solve(V, np.array([R(z) for z in zs]))
    np.testing.assert_allclose(coef[:5].real, [1, 1, 1 / 2, 1 / 6, 1 / 24], atol=1e-12)
    np.testing.assert_allclose(coef.imag, 0, atol=1e-12)
    # stage abscissae: apply to y' = 1 (t-integration) and read the stage times
    t, k = 0.0, 0.0
    for s in range(5):
        assert abs(t - c[s]) < 1e-14
        k = A[s] * k + 1.0
        t = t + B[s] * k
    assert abs(t - 1.0) < 1e-14
