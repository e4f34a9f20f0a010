"""Pins the oracle's geometry and BK3 operator (pmo_mesh.c, pmo_bk3.c)."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform


def test_affine_geometry_closed_form(orc):
    """Axis-aligned box cells: J = diag(h), G = JxW diag(1/h^2), sum JxW = volume."""
    m = orc.mesh(2, 3, 4, lo=(0.0, -1.0, 0.5), hi=(2.0, 0.5, 3.5))
    nq = 3
    G, J, _ = orc.geometry(m, nq, want_jxw=True)
    h = np.array([2.0 / 2, 1.5 / 3, 3.0 / 4])
    _, w = orc.gauss_legendre(nq)
    w3 = np.einsum("i,j,k->kji", w, w, w).ravel()
    G = G.reshape(-1, 6, nq ** 3)
    jxw = np.prod(h) * w3
    for comp, val in zip(range(6), [1 / h[0] ** 2, 0, 0, 1 / h[1] ** 2, 0, 1 / h[2] ** 2]):
        assert np.allclose(G[:, comp, :], jxw * val, atol=1e-15)
    assert abs(J.sum() - 2.0 * 1.5 * 3.0) < 1e-13


def test_sheared_cell_metric(orc):
    """Deformed (trilinear) cells: G = JxW J^-1 J^-T with J from finite
    differences of the physical quadrature points, and the volume sum JxW is
    independent of the quadrature order (det J is a polynomial of degree 2)."""
    m = orc.mesh(3, deform=1, amplitude=0.05)
    v2 = orc.geometry(m, 2, want_jxw=True)[1].sum()
    v6 = orc.geometry(m, 6, want_jxw=True)[1].sum()
    assert abs(v2 - v6) < 1e-13
    # single cell check against an explicit trilinear Jacobian
    cv = orc.cell_vertices(m).reshape(-1, 8, 3)[5]
    nq = 2
    G = orc.geometry(m, nq)[0].reshape(-1, 6, nq ** 3)[5]
    xq, wq = orc.gauss_legendre(nq)
    q = 0
    for r in range(nq):
        for s in range(nq):
            for t in range(nq):
                xi = np.array([xq[t], xq[s], xq[r]])
                J = np.zeros((3, 3))
                for v in range(8):
                    b = [(v >> k) & 1 for k in range(3)]
                    f = [xi[k] if b[k] else 1 - xi[k] for k in range(3)]
                    df = [1 if b[k] else -1 for k in range(3)]
                    dN = [df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]]
                    J += np.outer(cv[v], dN)
                Ji = np.linalg.inv(J)
                Gm = abs(np.linalg.det(J)) * wq[t] * wq[s] * wq[r] * Ji @ Ji.T
                ref = [Gm[0, 0], Gm[0, 1], Gm[0, 2], Gm[1, 1], Gm[1, 2], Gm[2, 2]]
                assert np.allclose(G[:, q], ref, atol=1e-14)
                q += 1


@pytest.mark.parametrize("p,nq,deform", [(1, 2, 0), (2, 3, 1), (2, 4, 0), (3, 4, 1),
                                         (4, 5, 1), (4, 6, 0)])
def test_sum_factorisation_equals_dense_element_matrices(orc, p, nq, deform):
    m = orc.mesh(2, deform=deform, amplitude=0.07)
    op = orc.LaplaceBK3(m, p, nq)
    K = op.element_matrices()
    N, nm3 = op.N, (p + 1) ** 3
    A = np.zeros((N, N))
    for c in range(op.C):
        idx = op.plain_map[c * nm3:(c + 1) * nm3]
        A[np.ix_(idx, idx)] += K[c]
    op.map = op.plain_map.copy()  # no constraints: pure cell loop
    x = seeded_uniform(N, 9)
    y = op.cells_add(x, np.zeros(N))
    assert rel_err(y, A @ x) < 1e-13
    assert np.allclose(A, A.T, atol=1e-13 * np.abs(A).max())
    # constant functions are in the kernel of the unconstrained operator
    assert np.abs(A @ np.ones(N)).max() < 1e-12 * np.abs(A).max()


def test_q1_stencil_closed_form(orc):
    """Q1 on a uniform grid: element matrix h/3 on the diagonal, 0 between
    edge neighbours, -h/12 between face- and body-diagonal vertices."""
    h = 0.25
    K = orc.LaplaceBK3(orc.mesh(4), 1, 2).element_matrices()[0]
    ref = np.zeros((8, 8))
    for a in range(8):
        for b in range(8):
            d = bin(a ^ b).count("1")
            ref[a, b] = {0: h / 3, 1: 0.0, 2: -h / 12, 3: -h / 12}[d]
    assert np.allclose(K, ref, atol=1e-15)


@pytest.mark.parametrize("p", [2, 3, 4])
def test_energy_of_polynomial(orc, p):
    """x^T A x = int |grad u|^2 for u = x^2 + y z on the unit cube (= 2)."""
    m = orc.mesh(3)
    op = orc.LaplaceBK3(m, p, p + 1)
    op.map = op.plain_map.copy()
    n1 = p * 3 + 1
    z = orc.gauss_lobatto(p + 1)
    coords = np.concatenate([[(c + zz) / 3 for zz in z[:-1]] for c in range(3)] + [[1.0]])
    X, Y, Z = np.meshgrid(coords, coords, coords, indexing="ij")
    u = (X ** 2 + Y * Z).transpose(2, 1, 0).ravel()  # x fastest
    assert len(u) == n1 ** 3
    y = op.cells_add(u, np.zeros(op.N))
    assert abs(u @ y - 2.0) < 1e-12


@pytest.mark.parametrize("p,nq,deform", [(1, 2, 0), (2, 3, 1), (4, 5, 0), (3, 5, 1)])
def test_diagonal_equals_dense(orc, p, nq, deform):
    m = orc.mesh(2, deform=deform, amplitude=0.05)
    op = orc.LaplaceBK3(m, p, nq)
    K = op.element_matrices()
    nm3 = (p + 1) ** 3
    d = np.zeros(op.N)
    for c in range(op.C):
        for i in range(nm3):
            g = op.map[c * nm3 + i]
            if g != orc.INVALID:
                d[g] += K[c, i, i]
    d[op.cons] = 1.0
    assert rel_err(op.diagonal(), d) < 1e-13


def test_vmult_framing_and_colours(orc):
    m = orc.mesh(4, deform=1, amplitude=0.05)
    op = orc.LaplaceBK3(m, 2, 3)
    x = seeded_uniform(op.N, 4)
    y1 = op.vmult(x, n_threads=1)
    y4 = op.vmult(x, n_threads=4)
    assert rel_err(y4, y1) < 1e-14
    assert np.array_equal(y1[op.cons], x[op.cons])
    # no two cells of one colour share a DoF
    nm3 = 27
    for col in range(8):
        cells = np.nonzero(op.colors == col)[0]
        dofs = np.concatenate([op.plain_map[c * nm3:(c + 1) * nm3] for c in cells])
        assert len(np.unique(dofs)) == len(dofs)
    # symmetric on the free DoFs
    z = seeded_uniform(op.N, 5)
    x[op.cons] = 0
    z[op.cons] = 0
    assert abs(op.vmult(x) @ z - x @ op.vmult(z)) < 1e-12 * abs(x @ op.vmult(z))


def test_gradient_matches_finite_difference_of_energy(orc):
    """d/dkappa_q of 0.5 u^T A(kappa) lam ... checked through the bilinear form:
    sum_q out_q dkappa_q / kappa_q = -u^T (A(kappa + dkappa) - A(kappa)) lam."""
    m = orc.mesh(2, deform=1, amplitude=0.05)
    op = orc.LaplaceBK3(m, 2, 3)
    G0 = op.G.copy()
    nq3 = 27
    kap = np.exp(seeded_uniform(op.C * nq3, 6, -0.3, 0.3))
    dk = seeded_uniform(op.C * nq3, 7) * 1e-3
    u, lam = seeded_uniform(op.N, 8), seeded_uniform(op.N, 10)
    u[op.cons] = 0
    lam[op.cons] = 0
    g = op.grad_kappa(kap, u, lam)
    op.G = orc.scale_G(G0, kap, 3)
    a0 = u @ op.vmult(lam)
    op.G = orc.scale_G(G0, kap + dk, 3)
    a1 = u @ op.vmult(lam)
    assert abs((g * dk / kap).sum() + (a1 - a0)) < 1e-12 * abs(a1 - a0) + 1e-15


@pytest.mark.parametrize("p,nq,deform", [(2, 3, 1), (3, 5, 0)])
def test_subdomain_vmults_equal_dense_assembly(orc, p, nq, deform):
    """vmult_neumann and vmult_interface_cell_range restated against a dense
    assembly of exact element matrices over the same cells, with the physical
    boundary rows and columns replaced by the identity."""
    from conftest import subdomain_sets
    m = orc.mesh(3, deform=deform, amplitude=0.06)
    op = orc.LaplaceBK3(m, p, nq)
    K = op.element_matrices()
    N, nm3 = op.N, (p + 1) ** 3
    phys, iface, icells = subdomain_sets(3, 3, 3, p)
    assert len(icells) == 9 and len(iface) == (3 * p - 1) ** 2

    def dense(cells):
        A = np.zeros((N, N))
        for c in cells:
            idx = op.plain_map[c * nm3:(c + 1) * nm3]
            A[np.ix_(idx, idx)] += K[c]
        A[phys, :] = 0.0
        A[:, phys] = 0.0
        A[phys, phys] = 1.0
        return A

    x = seeded_uniform(N, 11)
    y_neu = op.vmult_subdomain(x, phys)
    assert rel_err(y_neu, dense(range(op.C)) @ x) < 1e-13
    y_int = op.vmult_subdomain(x, phys, icells)
    assert rel_err(y_int, dense(icells) @ x) < 1e-13
    # interface cells + the other cells = every cell, on the free rows
    others = np.setdiff1d(np.arange(op.C), icells).astype(np.uint32)
    y_oth = op.vmult_subdomain(x, phys, others)
    free = np.setdiff1d(np.arange(N), phys)
    assert rel_err((y_int + y_oth)[free], y_neu[free]) < 1e-13
    # with every boundary DoF physical it is the operator's vmult
    assert rel_err(op.vmult_subdomain(x, op.cons), op.vmult(x)) < 1e-14
    # without physical DoFs constants are in its kernel
    zero = op.vmult_subdomain(np.ones(N), np.zeros(0, np.uint32))
    assert np.abs(zero).max() < 1e-12 * np.abs(y_neu).max()
