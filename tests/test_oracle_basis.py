"""Pins the oracle's 1D data (pmo_basis.c) with closed forms and library routines."""
import numpy as np
import pytest
from numpy.polynomial import legendre as npleg


@pytest.mark.parametrize("n", range(1, 11))
def test_gauss_legendre_matches_numpy(orc, n):
    x, w = orc.gauss_legendre(n)
    t, wt = npleg.leggauss(n)
    assert np.allclose(x, 0.5 * (t + 1), atol=1e-15, rtol=0)
    assert np.allclose(w, 0.5 * wt, atol=1e-15, rtol=0)
    for k in range(2 * n):  # exact up to degree 2n-1
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) < 1e-14


@pytest.mark.parametrize("n", range(2, 10))
def test_gauss_lobatto_are_roots_of_legendre_derivative(orc, n):
    z = orc.gauss_lobatto(n)
    assert z[0] == 0.0 and z[-1] == 1.0
    c = np.zeros(n)
    c[-1] = 1.0  # P_{n-1}
    roots = np.sort(npleg.legroots(npleg.legder(c))) if n > 2 else np.array([])
    assert np.allclose(z[1:-1], 0.5 * (roots + 1), atol=1e-14, rtol=0)
    assert np.allclose(z, 1 - z[::-1], atol=1e-15)


@pytest.mark.parametrize("m", [2, 3, 5, 7])
def test_lagrange_basis(orc, m):
    z = orc.gauss_lobatto(m)
    for j, zj in enumerate(z):  # Kronecker property
        v, _ = orc.lagrange(z, zj)
        assert np.allclose(v, np.eye(m)[j], atol=1e-14)
    rng = np.random.default_rng(3)
    coef = rng.standard_normal(m)  # random polynomial of degree m-1
    poly = np.polynomial.Polynomial(coef)
    for x in rng.uniform(0, 1, 5):
        v, d = orc.lagrange(z, x)
        assert abs(v.sum() - 1) < 1e-13 and abs(d.sum()) < 1e-11
        assert abs(v @ poly(z) - poly(x)) < 1e-12
        assert abs(d @ poly(z) - poly.deriv()(x)) < 1e-10


@pytest.mark.parametrize("p,nq", [(1, 2), (2, 3), (4, 5), (4, 6), (6, 7), (6, 8)])
def test_shape_matrices_reproduce_polynomials(orc, p, nq):
    """S interpolates Q_p data to the Gauss points, D differentiates exactly on
    the Gauss points, and S then D equals the exact basis derivative."""
    S, D = orc.shape_matrices(p, nq)
    S, D = S.reshape(p + 1, nq), D.reshape(nq, nq)
    z = orc.gauss_lobatto(p + 1)
    xq, _ = orc.gauss_legendre(nq)
    poly = np.polynomial.Polynomial(np.random.default_rng(p).standard_normal(p + 1))
    uq = poly(z) @ S
    assert np.allclose(uq, poly(xq), atol=1e-12)
    assert np.allclose(uq @ D, poly.deriv()(xq), atol=1e-10)
    assert np.allclose(D.sum(0), 0, atol=1e-11)  # D * 1 = 0
    Sd = orc.shape_gradients(p, nq).reshape(p + 1, nq)
    assert np.allclose(S @ D, Sd, atol=1e-11)


def test_tridiagonal_eigenvalues(orc):
    rng = np.random.default_rng(5)
    d, o = rng.uniform(1, 3, 12), rng.uniform(-1, 1, 11)
    T = np.diag(d) + np.diag(o, 1) + np.diag(o, -1)
    assert np.allclose(orc.tridiag_eigs(d, o), np.linalg.eigvalsh(T), atol=1e-12)
