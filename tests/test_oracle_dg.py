"""Pins the oracle's SIP-DG operator (pmo_dg.c) without the reference:
symmetry and coercivity, reduction to the continuous operator on continuous
functions, Galerkin consistency on sheared cells (decides J^{-1}n against
J^{-T}n), closed-form penalties, and the DG<->CG transfer."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform


def dense(op):
    cols = []
    for j in range(op.N):
        e = np.zeros(op.N)
        e[j] = 1.0
        cols.append(op.vmult(e))
    return np.array(cols).T


@pytest.mark.parametrize("metric", [0, 1])
def test_dg_symmetric_positive_definite(orc, metric):
    m = orc.mesh(2, deform=1, amplitude=0.05)
    op = orc.LaplaceDG(m, 2, metric=metric)
    A = dense(op)
    assert rel_err(A, A.T) < 1e-13
    ev = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert ev.min() > 1e-3 * ev.max() * 1e-3


@pytest.mark.parametrize("p", [1, 2, 3])
def test_dg_equals_cg_operator_on_continuous_functions(orc, p):
    """For continuous u, v vanishing on the boundary every face term drops:
    P^T A_DG P = A_CG on the free DoFs (P = DG<-CG injection)."""
    m = orc.mesh(3, 2, 2, deform=1, amplitude=0.05)
    dg = orc.LaplaceDG(m, p)
    cg = orc.LaplaceBK3(m, p)
    cg_map = orc.dofmap(m, p, True)
    x = seeded_uniform(cg.N, 11)
    x[cg.cons] = 0.0
    u = orc.dg_cg_prolongate(p, cg_map, x, np.zeros(dg.N))
    y = orc.dg_cg_restrict(p, cg_map, dg.vmult(u), np.zeros(cg.N))
    ref = cg.vmult(x)
    free = np.setdiff1d(np.arange(cg.N), cg.cons)
    assert rel_err(y[free], ref[free]) < 1e-12


def _sheared_problem(orc, n, p, a):
    """u = prod_d g(l_d(x)), g(t) = t(1-t), l = B x the inverse shear, on the
    sheared unit cube; returns (mesh, nodal u, f = -Laplace u as a callable)."""
    A = np.array([[1.0, a, 0.0], [0.0, 1.0, a], [a, 0.0, 1.0]])
    B = np.linalg.inv(A)
    M = B @ B.T

    def lvals(X):
        return X @ B.T

    def u_of(X):
        L = lvals(X)
        return np.prod(L * (1 - L), axis=-1)

    def f_of(X):
        L = lvals(X)
        g = L * (1 - L)
        dg = 1 - 2 * L
        lap = np.zeros(X.shape[:-1])
        for d in range(3):
            for e in range(3):
                if d == e:
                    t = -2.0 * np.prod([g[..., k] for k in range(3) if k != d], axis=0)
                else:
                    k = 3 - d - e
                    t = dg[..., d] * dg[..., e] * g[..., k]
                lap = lap + M[d, e] * t
        return -lap

    m = orc.mesh(n, deform=2, amplitude=a)
    return m, A, u_of, f_of


def test_dg_galerkin_consistency_on_sheared_cells(orc):
    """A smooth u in the discrete space with u = 0 on the boundary satisfies
    a(u, phi_i) = int (-Laplace u) phi_i exactly.  On parallelepiped cells
    (non-symmetric J) this holds with d_n = (J^{-1} n).grad_xi and fails with
    (J^{-T} n).grad_xi."""
    p, n, a, nq = 2, 2, 0.3, 6
    m, A, u_of, f_of = _sheared_problem(orc, n, p, a)
    nm = p + 1
    gll = orc.gauss_lobatto(nm)
    xq, wq = orc.gauss_legendre(nq)
    cv = orc.cell_vertices(m).reshape(-1, 8, 3)
    C = cv.shape[0]
    h = 1.0 / n
    detJ = abs(np.linalg.det(A)) * h ** 3
    u = np.zeros(C * nm ** 3)
    b = np.zeros(C * nm ** 3)
    Sq = np.array([orc.lagrange(gll, x)[0] for x in xq])  # [q][i]
    for c in range(C):
        x0 = cv[c, 0]
        Jc = A * h  # affine cell map x = x0 + Jc xi
        xi_n = np.stack(np.meshgrid(gll, gll, gll, indexing="ij"), -1)[..., ::-1]  # [k][j][i][xyz]
        u[c * nm ** 3:(c + 1) * nm ** 3] = u_of(x0 + xi_n.reshape(-1, 3) @ Jc.T)
        xi_q = np.stack(np.meshgrid(xq, xq, xq, indexing="ij"), -1)[..., ::-1].reshape(-1, 3)
        fq = f_of(x0 + xi_q @ Jc.T).reshape(nq, nq, nq)  # [r][s][t] = z,y,x
        w3 = np.einsum("r,s,t->rst", wq, wq, wq) * detJ
        bc = np.einsum("rst,tl,sj,rk->kjl", fq * w3, Sq, Sq, Sq)
        b[c * nm ** 3:(c + 1) * nm ** 3] = bc.ravel()
    good = orc.LaplaceDG(m, p, nq=p + 1, metric=0).vmult(u)
    bad = orc.LaplaceDG(m, p, nq=p + 1, metric=1).vmult(u)
    assert rel_err(good, b) < 1e-12
    assert rel_err(bad, b) > 1e-3


def test_dg_metric_variants_agree_on_axis_aligned_cells(orc):
    m = orc.mesh(2, 3, 2, lo=(0.0, 0.0, -1.0), hi=(1.0, 2.0, 0.5))
    x = seeded_uniform(12 * 27, 3)
    a0 = orc.LaplaceDG(m, 2, metric=0).vmult(x)
    a1 = orc.LaplaceDG(m, 2, metric=1).vmult(x)
    assert rel_err(a0, a1) < 1e-14


@pytest.mark.parametrize("kind", [0, 1])
def test_dg_penalty_closed_form(orc, kind):
    """Uniform cells of size h: sigma = p(p+1)/h (kind 0) or 2(p+1)^2/h
    (kind 1); for u = 1 on one cell, a(u,u) = sum over its faces of
    sigma |F| (inner) or 2 sigma |F| (boundary: u+ = -u-)."""
    p, n, L = 3, 3, 1.5
    h = L / n
    m = orc.mesh(n, hi=(L, L, L))
    op = orc.LaplaceDG(m, p, penalty_kind=kind)
    sig = p * (p + 1) / h if kind == 0 else 2 * (p + 1) ** 2 / h
    assert np.allclose(op.sigma, sig, rtol=1e-13)
    nm3 = (p + 1) ** 3
    for c, n_bdry in [(13, 0), (0, 3), (4, 1)]:
        x = np.zeros(op.N)
        x[c * nm3:(c + 1) * nm3] = 1.0
        e = x @ op.vmult(x)
        assert abs(e - sig * h * h * ((6 - n_bdry) + 2 * n_bdry)) < 1e-11 * e


def test_dg_neumann_annihilates_constants(orc):
    m = orc.mesh(2, deform=1, amplitude=0.05)
    op = orc.LaplaceDG(m, 2, neumann_mask=0x3F)
    assert np.abs(op.vmult(np.ones(op.N))).max() < 1e-12


def test_dg_cg_transfer_adjoint_and_injection(orc):
    p = 3
    m = orc.mesh(2, 3, 2)
    cg_map = orc.dofmap(m, p, True)
    Ncg, Ndg = orc.n_dofs(m, p), orc.n_cells(m) * (p + 1) ** 3
    x, y = seeded_uniform(Ncg, 1), seeded_uniform(Ndg, 2)
    cons = orc.boundary_dofs(m, p)
    Px = orc.dg_cg_prolongate(p, cg_map, x, np.zeros(Ndg))
    Ry = orc.dg_cg_restrict(p, cg_map, y, np.zeros(Ncg))
    assert abs(Px @ y - x @ Ry) < 1e-12 * np.abs(Px @ y)
    assert np.all(Ry[cons] == 0)
    # every DG DoF takes the value of the CG DoF at its support point
    plain = orc.dofmap(m, p, False)
    x[cons] = 0.0
    assert np.array_equal(orc.dg_cg_prolongate(p, cg_map, x, np.zeros(Ndg)), x[plain])
