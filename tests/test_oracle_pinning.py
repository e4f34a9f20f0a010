"""Pins of the oracle functions that no reference test fixes, by what the
mathematics fixes:

- the Chebyshev smoother's eigenvalue estimate (deal.II
  internal::estimate_eigenvalues as configured at
  tests/poisson_cube_bk3/program.cc:474-515: CG on P^{-1}A from the start
  vector (i mod 11) minus its mean, IterationNumberControl(k, 1e-10), Ritz
  values of the Lanczos tridiagonal, 1.2 safety factor on the largest) is a
  Rayleigh-Ritz projection: its Ritz values are the eigenvalues of the pencil
  (V^T A V, V^T D V) on the Krylov space V = K_k(D^{-1}A, D^{-1}b).  A wrong
  iteration count, start vector, preconditioner or safety factor moves them;
- the automatic degree (range < 1) is the smallest k with
  1 / T_k(theta / delta) <= range, the Chebyshev iteration's error bound;
- the SIP penalty's h = |K| / |F| on non-planar faces (trilinear cells of the
  sinusoidal deformation) against a 20 x 20 Gauss rule of |d_s x  x  d_t x|;
- the right-hand side and the L2 error against closed-form integrals.
"""
import numpy as np
import pytest
import scipy.linalg


def _dense_level(ref, level):
    n = ref.n_dofs(level)
    A = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = ref.level_vmult(level, e)
        e[j] = 0.0
    return A


def _start_vector(n):
    b = (np.arange(n) % 11).astype(np.float64)
    return b - b.mean()


def _ritz_values(A, d, b, k):
    """Eigenvalues of V^T A V y = theta V^T D V y, V a D-orthonormal basis of
    K_k(D^{-1}A, D^{-1}b) (Arnoldi with two Gram-Schmidt passes)."""
    n = len(b)
    V = np.zeros((n, k))
    v = b / d
    for j in range(k):
        for _ in range(2):
            v = v - V[:, :j] @ (V[:, :j].T @ (d * v))
        nrm = np.sqrt(v @ (d * v))
        V[:, j] = v / nrm
        v = (A @ V[:, j]) / d
    H = V.T @ A @ V
    M = V.T @ (d[:, None] * V)
    return scipy.linalg.eigh(0.5 * (H + H.T), 0.5 * (M + M.T), eigvals_only=True)


def _cg_iterations(A, d, b, n_it):
    """Iterations the estimate runs: k <= n_it, stopped once ||r|| <= 1e-10."""
    x, r = np.zeros_like(b), b.copy()
    z = r / d
    p = z.copy()
    rho = r @ z
    k = 0
    while k < n_it and np.sqrt(r @ r) > 1e-10:
        q = A @ p
        a = rho / (p @ q)
        x += a * p
        r -= a * q
        k += 1
        z = r / d
        rho_new = r @ z
        p = z + rho_new / rho * p
        rho = rho_new
    return k


@pytest.mark.parametrize("p,deform,n_it,n", [(2, 0, 10, 2), (3, 1, 10, 2), (2, 1, 6, 2),
                                           (1, 1, 25, 4)])
def test_chebyshev_eigenvalue_estimate_is_rayleigh_ritz(orc, p, deform, n_it, n):
    meshes = [orc.mesh(n // 2, deform=deform, amplitude=0.05),
              orc.mesh(n, deform=deform, amplitude=0.05)]
    ref = orc.Multigrid(meshes, [p, p], eig_cg_n_iterations=n_it)
    A = _dense_level(ref, 1)
    d = np.diag(A).copy()
    b = _start_vector(len(d))
    k = _cg_iterations(A, d, b, n_it)
    assert k >= 3
    theta = _ritz_values(A, d, b, k)
    max_ev, min_ev, _ = ref.smoother_info(1)
    assert abs(min_ev - theta[0]) <= 1e-9 * theta[-1]
    assert abs(max_ev - 1.2 * theta[-1]) <= 1e-10 * theta[-1]
    if n_it > 10:
        return  # Krylov space nearly exhausted: every start vector gives these values
    # at the smoothers' 10 steps a plausible slip is caught: one CG step fewer,
    # or the start vector without its mean removed
    alt = _ritz_values(A, d, b, k - 1)
    assert abs(max_ev - 1.2 * alt[-1]) > 1e-8 * theta[-1] or abs(min_ev - alt[0]) > 1e-8
    alt = _ritz_values(A, d, (np.arange(len(d)) % 11).astype(np.float64), k)
    assert abs(max_ev - 1.2 * alt[-1]) > 1e-8 * theta[-1] or abs(min_ev - alt[0]) > 1e-8


def _cheb_T(k, x):
    return np.cosh(k * np.arccosh(x))


@pytest.mark.parametrize("n,p", [(2, 2), (2, 1), (1, 3)])
def test_chebyshev_auto_degree_is_the_smallest_that_meets_the_bound(orc, n, p):
    """Coarse smoother, degree automatic for relative accuracy 1e-3 with
    eig_cg_n_iterations = m() (CG then stops at ||r|| <= 1e-10): its Ritz
    values are the Rayleigh-Ritz values of that Krylov space, and the degree
    is the smallest k
    whose Chebyshev bound 1/T_k(theta/delta) reaches 1e-3 on
    [alpha, lambda_max], alpha = min(0.9 lambda_max, lambda_min)."""
    m = orc.mesh(n, deform=1, amplitude=0.05)
    ref = orc.Multigrid([m], [p], coarse_kind=0)
    max_ev, min_ev, deg = ref.smoother_info(0)
    A = _dense_level(ref, 0)
    d = np.diag(A).copy()
    b = _start_vector(len(d))
    k = _cg_iterations(A, d, b, len(d))
    ev = _ritz_values(A, d, b, k)
    assert abs(max_ev - 1.2 * ev[-1]) <= 1e-9 * ev[-1]
    assert abs(min_ev - ev[0]) <= 1e-8 * ev[-1]
    alpha = min(0.9 * max_ev, min_ev)
    theta, delta = 0.5 * (max_ev + alpha), 0.5 * (max_ev - alpha)
    assert 1.0 / _cheb_T(deg, theta / delta) <= 1e-3
    assert 1.0 / _cheb_T(deg - 1, theta / delta) > 1e-3


def _trilinear_jacobian(cv, xi):
    """J[d][k] = d x_d / d xi_k of the trilinear map of the 8 vertices cv[v][d]."""
    J = np.zeros((3, 3))
    for v in range(8):
        bits = [(v >> a) & 1 for a in range(3)]
        f = [xi[a] if bits[a] else 1.0 - xi[a] for a in range(3)]
        s = [1.0 if bits[a] else -1.0 for a in range(3)]
        dN = [s[0] * f[1] * f[2], f[0] * s[1] * f[2], f[0] * f[1] * s[2]]
        for k in range(3):
            J[:, k] += dN[k] * cv[v]
    return J


@pytest.mark.parametrize("kind", [0, 1])
def test_dg_penalty_face_measure_on_nonplanar_faces(orc, kind):
    """sigma = fac (1/h- + 1/h+), h = |K|/|F|, on the sinusoidally deformed
    mesh whose faces are not planar.  The oracle integrates the area element
    |d_s x  x  d_t x| with 4 x 4 Gauss points; it agrees with a 20 x 20 Gauss
    reference to 1e-6 (the 4-point rule's error on these faces is ~2e-7),
    and a 3 x 3 rule would not (~1e-5), so the test sees the rule.  deal.II's
    face->measure() is a different approximation of the same area; the
    difference is of this order and documented, not pinned."""
    p, n = 3, 3
    m = orc.mesh(n, deform=1, amplitude=0.08)
    op = orc.LaplaceDG(m, p, penalty_kind=kind)
    cv = orc.cell_vertices(m).reshape(-1, 8, 3)

    def rule(k):
        g, w = np.polynomial.legendre.leggauss(k)
        return 0.5 * (g + 1.0), 0.5 * w

    def measures(c, k):
        g, w = rule(k)
        gv, wv = rule(4)  # det J of a trilinear map: exact with 4 points
        vol = 0.0
        for a in range(4):
            for b_ in range(4):
                for c_ in range(4):
                    vol += abs(np.linalg.det(_trilinear_jacobian(cv[c], (gv[a], gv[b_], gv[c_])))) \
                        * wv[a] * wv[b_] * wv[c_]
        area = np.zeros(6)
        for f in range(6):
            dn = f // 2
            t1, t2 = (dn + 1) % 3, (dn + 2) % 3
            for a in range(k):
                for b_ in range(k):
                    xi = [0.0, 0.0, 0.0]
                    xi[dn], xi[t1], xi[t2] = float(f % 2), g[a], g[b_]
                    J = _trilinear_jacobian(cv[c], xi)
                    area[f] += np.linalg.norm(np.cross(J[:, t1], J[:, t2])) * w[a] * w[b_]
        return vol / area

    def sigma(k):
        h = np.array([measures(c, k) for c in range(n ** 3)])
        fac = 0.5 * p * (p + 1.0) if kind == 0 else (p + 1.0) ** 2
        sig = np.zeros((n ** 3, 6))
        for c in range(n ** 3):
            cx, cy, cz = c % n, (c // n) % n, c // (n * n)
            for f in range(6):
                cc = [cx, cy, cz]
                cc[f // 2] += 1 if f % 2 else -1
                hp = h[cc[0] + n * (cc[1] + n * cc[2]), f ^ 1] if 0 <= cc[f // 2] < n else h[c, f]
                sig[c, f] = fac * (1.0 / h[c, f] + 1.0 / hp)
        return sig

    ref = sigma(20)
    got = op.sigma.reshape(-1, 6)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-6
    assert np.abs(sigma(3) - ref).max() / np.abs(ref).max() > 1e-6


def _bubble_nodal(orc, n, p):
    """Nodal values of u = x(1-x) y(1-y) z(1-z) on the Q_p support points of
    the unit cube with n^3 cells (lexicographic global numbering)."""
    xg = orc.gauss_lobatto(p + 1)
    N1 = p * n + 1
    I = np.arange(N1)
    c = np.minimum(I // p, n - 1)
    x1 = (c + xg[I - c * p]) / n
    f1 = x1 * (1.0 - x1)
    return (f1[None, None, :] * f1[None, :, None] * f1[:, None, None]).ravel()


@pytest.mark.parametrize("n,p", [(2, 2), (3, 3), (2, 4)])
def test_rhs_and_l2_error_closed_forms(orc, n, p):
    """For the bubble u = x(1-x)y(1-y)z(1-z), exactly represented at p >= 2
    and zero on the Dirichlet boundary:
      b(f=1) . u_I = int u = 1/216 (exact quadrature);
      b(f=3 pi^2 sin sin sin) . u_I = 3 pi^2 (4/pi^3)^3;
      ||u_I - 0|| = (1/30)^(3/2);  ||0 - sin sin sin|| = (1/2)^(3/2)."""
    m = orc.mesh(n)
    op = orc.LaplaceBK3(m, p)
    u = _bubble_nodal(orc, n, p)
    assert abs(op.rhs(0) @ u - 1.0 / 216.0) <= 1e-14
    exact = 3.0 * np.pi ** 2 * (4.0 / np.pi ** 3) ** 3
    assert abs(op.rhs(1, nq=p + 3) @ u - exact) <= 1e-6 * exact
    assert abs(op.l2_error(u, 0) - (1.0 / 30.0) ** 1.5) <= 1e-14
    assert abs(op.l2_error(np.zeros(op.N), 1, nq=p + 4) - 0.5 ** 1.5) <= 1e-7
