"""Pins the oracle's Chebyshev smoother, PCG and V-cycle (pmo_solver.c):
the smoother equals the textbook Chebyshev residual polynomial evaluated by a
dense eigendecomposition, Ritz values lie inside the spectrum, PCG reproduces
the dense solve, and MG-CG converges at the discretisation's O(h^{p+1}) rate."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform


def _dense(orc, mg, level, n):
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = mg.level_vmult(level, e)
    return A


def _cheb_T(k, x):
    return np.cos(k * np.arccos(x)) if np.all(np.abs(x) <= 1) else np.where(
        np.abs(x) <= 1, np.cos(k * np.arccos(np.clip(x, -1, 1))),
        np.cosh(k * np.arccosh(np.maximum(np.abs(x), 1))) * np.sign(x) ** k)


@pytest.mark.parametrize("p,degree", [(1, 3), (2, 3), (2, 5)])
def test_chebyshev_is_the_textbook_polynomial(orc, p, degree):
    mg = orc.Multigrid([orc.mesh(2), orc.mesh(3)], [p, p], cheby_degree=degree)
    n = mg.n_dofs(1)
    A = _dense(orc, mg, 1, n)
    d = np.diag(A).copy()
    mx, mn, deg = mg.smoother_info(1)
    assert deg == degree
    alpha = mx / 15.0
    theta, delta = 0.5 * (mx + alpha), 0.5 * (mx - alpha)
    Dm = np.diag(1 / np.sqrt(d))
    lam, V = np.linalg.eigh(Dm @ A @ Dm)  # spectrum of P^{-1} A
    assert mx / 1.2 <= lam.max() * (1 + 1e-12) and mn >= lam.min() * (1 - 1e-12)
    # residual polynomial R(t) = T_k((theta - t)/delta) / T_k(theta/delta)
    R = _cheb_T(degree, (theta - lam) / delta) / _cheb_T(degree, np.array(theta / delta))
    b = seeded_uniform(n, 3)
    x = mg.smooth(1, np.zeros(n), b, 0)
    # x = (I - R(P^-1 A)) A^-1 b
    ref = Dm @ V @ np.diag((1 - R) / lam) @ V.T @ Dm @ b
    assert rel_err(x, ref) < 1e-10
    # step from a nonzero start: error e -> R(P^-1 A) e
    x0 = seeded_uniform(n, 4)
    xs = mg.smooth(1, x0.copy(), b, 1)
    xstar = np.linalg.solve(A, b)
    e_ref = Dm @ V @ np.diag(R) @ V.T @ np.diag(np.sqrt(d)) @ (x0 - xstar)
    assert rel_err(xs - xstar, e_ref) < 1e-9


def test_pcg_reproduces_dense_solve(orc):
    mg = orc.Multigrid([orc.mesh(3, deform=1, amplitude=0.05)], [2])
    n = mg.n_dofs(0)
    A = _dense(orc, mg, 0, n)
    b = seeded_uniform(n, 5)
    x, its, r0, r = mg.pcg(b, level=0, precond=1, rel_tol=1e-12, max_it=1000)
    assert r <= 1e-12 * r0 and its < n
    assert rel_err(x, np.linalg.solve(A, b)) < 1e-10
    x0, its0, _, _ = mg.pcg(b, level=0, precond=0, rel_tol=1e-12, max_it=1000)
    assert rel_err(x0, x) < 1e-10


@pytest.mark.parametrize("p", [1, 2, 3])
def test_mgcg_converges_at_the_expected_rate(orc, p):
    errs, its = [], []
    for n in (4, 8):
        meshes = [orc.mesh(n // 2), orc.mesh(n)]
        mg = orc.Multigrid(meshes, [p, p])
        op = orc.LaplaceBK3(meshes[-1], p)
        b = op.rhs(1)
        x, it, r0, r = mg.pcg(b, precond=2, rel_tol=1e-12)
        assert r <= 1e-12 * r0
        its.append(it)
        errs.append(op.l2_error(x, 1))
    rate = np.log2(errs[0] / errs[1])
    assert rate > p + 1 - 0.35, (errs, rate)
    assert max(its) < 20


def test_vcycle_is_a_symmetric_preconditioner(orc):
    """With an exact coarse solve the V-cycle (pre = post Chebyshev) is symmetric."""
    mg = orc.Multigrid([orc.mesh(2), orc.mesh(4)], [2, 2], coarse_tol=1e-15)
    n = mg.n_dofs(1)
    cons = orc.boundary_dofs(orc.mesh(4), 2)
    r1, r2 = seeded_uniform(n, 6), seeded_uniform(n, 7)
    r1[cons] = 0
    r2[cons] = 0
    a, b = mg.vmult(r1) @ r2, r1 @ mg.vmult(r2)
    assert abs(a - b) < 1e-9 * abs(a)
