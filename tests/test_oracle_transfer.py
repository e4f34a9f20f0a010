"""Pins the oracle's h/p transfers (pmo_transfer.c) with closed forms:
exact prolongation of polynomials of the coarse degree, partition of unity,
restriction = prolongation transpose, and multiplicity weights."""
import numpy as np
import pytest

from conftest import seeded_uniform


def _grid_coords(n, p):
    z = np.array(__import__("oracle").gauss_lobatto(p + 1))
    return np.concatenate([[(c + zz) / n for zz in z[:-1]] for c in range(n)] + [[1.0]])


def _nodal(n, p, f):
    c = _grid_coords(n, p)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    return f(X, Y, Z).transpose(2, 1, 0).ravel()


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_h_matrix_closed_form(orc, p):
    M = orc.h_matrix(p).reshape(p + 1, 2 * p + 1)
    assert np.allclose(M.sum(0), 1.0, atol=1e-14)
    z = orc.gauss_lobatto(p + 1)
    fine = np.concatenate([0.5 * z[:-1], 0.5 * (1 + z)])
    poly = np.polynomial.Polynomial(np.random.default_rng(p).standard_normal(p + 1))
    assert np.allclose(poly(z) @ M, poly(fine), atol=1e-12)


@pytest.mark.parametrize("pc,pf", [(1, 2), (2, 4), (3, 6), (1, 3)])
def test_p_matrix_closed_form(orc, pc, pf):
    M = orc.p_matrix(pc, pf).reshape(pc + 1, pf + 1)
    zc, zf = orc.gauss_lobatto(pc + 1), orc.gauss_lobatto(pf + 1)
    poly = np.polynomial.Polynomial(np.random.default_rng(pf).standard_normal(pc + 1))
    assert np.allclose(poly(zc) @ M, poly(zf), atol=1e-12)


def _poly(X, Y, Z):
    return (X * (1 - X)) * (Y * (1 - Y)) * (Z * (1 - Z)) * (1 + X + 2 * Y - Z)


@pytest.mark.parametrize("p", [1, 2, 4])
def test_h_prolongation_is_exact_for_coarse_polynomials(orc, p):
    """u vanishing on the boundary and of degree <= p per direction is
    prolongated to its fine nodal values (weights average identical values)."""
    n = 2
    if p < 4:
        f = lambda X, Y, Z: (X * (1 - X)) * (Y * (1 - Y)) * (Z * (1 - Z)) if p >= 2 else \
            np.minimum.reduce([X, 1 - X, Y, 1 - Y, Z, 1 - Z])
    else:
        f = _poly
    if p == 1:
        # piecewise-trilinear interpolant of a hat: compare with the coarse Q1 field itself
        uc = _nodal(n, 1, lambda X, Y, Z: np.minimum.reduce([X, 1 - X, Y, 1 - Y, Z, 1 - Z]))
        t = orc.Transfer(orc.mesh(n), 1, kind="h")
        uf = t.prolongate_add(uc, np.zeros((2 * n + 1) ** 3))
        c = _grid_coords(2 * n, 1)
        X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
        # trilinear interpolation of the coarse nodal values at the fine nodes
        from scipy.interpolate import RegularGridInterpolator
        cc = _grid_coords(n, 1)
        interp = RegularGridInterpolator((cc, cc, cc), uc.reshape(n + 1, n + 1, n + 1).transpose(2, 1, 0))
        ref = interp(np.stack([X, Y, Z], -1)).transpose(2, 1, 0).ravel()
        inner = _interior_mask(2 * n, 1)
        assert np.allclose(uf[inner], ref[inner], atol=1e-14)
        return
    uc = _nodal(n, p, f)
    t = orc.Transfer(orc.mesh(n), p, kind="h")
    uf = t.prolongate_add(uc, np.zeros((2 * n * p + 1) ** 3))
    ref = _nodal(2 * n, p, f)
    inner = _interior_mask(2 * n, p)
    assert np.allclose(uf[inner], ref[inner], atol=1e-13)
    assert np.allclose(uf[~inner], 0.0)


def _interior_mask(n, p):
    N1 = n * p + 1
    I = np.arange(N1)
    b = (I == 0) | (I == N1 - 1)
    B = b[None, None, :] | b[None, :, None] | b[:, None, None]
    return ~B.ravel()


@pytest.mark.parametrize("pc,pf", [(1, 2), (2, 4), (3, 6)])
def test_p_prolongation_is_exact(orc, pc, pf):
    n = 2
    f = lambda X, Y, Z: (X * (1 - X)) * (Y * (1 - Y)) * (Z * (1 - Z)) if pc >= 2 else \
        np.sin(X * 0)  # pc = 1 -> only the zero field vanishes on the boundary at n=2 nodes
    uc = _nodal(n, pc, f)
    t = orc.Transfer(orc.mesh(n), pc, pf, kind="p")
    uf = t.prolongate_add(uc, np.zeros((n * pf + 1) ** 3))
    assert np.allclose(uf, _nodal(n, pf, f), atol=1e-13)


@pytest.mark.parametrize("kind,pc,pf", [("h", 1, None), ("h", 3, None), ("p", 2, 5), ("p", 1, 4)])
def test_restriction_is_prolongation_transpose(orc, kind, pc, pf):
    m = orc.mesh(2, 3, 2)
    t = orc.Transfer(m, pc, pf, kind=kind)
    if kind == "h":
        nc, nf = orc.n_dofs(m, pc), orc.n_dofs(orc.mesh(4, 6, 4), pc)
    else:
        nc, nf = orc.n_dofs(m, pc), orc.n_dofs(m, pf)
    x, y = seeded_uniform(nc, 1), seeded_uniform(nf, 2)
    Px = t.prolongate_add(x, np.zeros(nf))
    Ry = t.restrict_add(y, np.zeros(nc))
    assert abs(Px @ y - x @ Ry) < 1e-12 * abs(Px @ y)


def test_weights_are_inverse_multiplicity(orc):
    t = orc.Transfer(orc.mesh(3), 2, kind="h")
    nf = orc.n_dofs(orc.mesh(6), 2)
    acc = np.zeros(nf)
    np.add.at(acc, t.map_f, t.w)
    inner = _interior_mask(6, 2)
    assert np.allclose(acc[inner], 1.0) and np.allclose(acc[~inner], 0.0)
