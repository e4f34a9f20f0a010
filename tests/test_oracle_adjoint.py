"""Pins the adjoint gradient of the variable-coefficient problem in the oracle
end to end: J(g) = 1/2 |u(g) - u_d|^2 with A(exp(g)) u = b, gradient
dJ/dg_q = -kappa_q (grad u)^T G_q (grad lambda), A lambda = u - u_d, checked
against central finite differences of J along random directions."""
import numpy as np

from conftest import seeded_uniform


def fourier_g(X, seed, n_modes=16, kmax=4):
    """g(x) = sum_m a_m cos(2 pi k_m . x + phi_m), integer wave vectors with
    max-norm 1..kmax, a_m ~ N(0, 0.5/|k_m|^2) (seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = np.zeros(X.shape[0])
    for _ in range(n_modes):
        while True:
            k = rng.integers(-kmax, kmax + 1, 3)
            if np.abs(k).max() >= 1:
                break
        a = rng.normal(0.0, np.sqrt(0.5 / float(k @ k)))
        phi = rng.uniform(0, 2 * np.pi)
        g += a * np.cos(2 * np.pi * (X @ k) + phi)
    return g


def test_adjoint_gradient_matches_central_differences(orc):
    n, p, nq = 3, 2, 3
    m = orc.mesh(n)
    G0, _, X = orc.geometry(m, nq, want_x=True)
    g0 = fourier_g(X.reshape(-1, 3), 5)
    base = orc.LaplaceBK3(m, p, nq)
    free = np.setdiff1d(np.arange(base.N), base.cons)
    b = base.rhs(0)

    def op_for(g):
        return orc.LaplaceBK3(m, p, nq, G=orc.scale_G(G0, np.exp(g), nq))

    def dense(op):
        A = np.zeros((base.N, base.N))
        for j in range(base.N):
            e = np.zeros(base.N)
            e[j] = 1.0
            A[:, j] = op.vmult(e)
        return A

    def solve(g, rhs):
        return np.linalg.solve(dense(op_for(g)), rhs)

    u_d = solve(np.zeros_like(g0), b)
    J = lambda g: 0.5 * np.sum((solve(g, b) - u_d) ** 2)
    u = solve(g0, b)
    lam = solve(g0, u - u_d)
    assert np.all(u[base.cons] == 0) and np.all(lam[base.cons] == 0)
    op = op_for(np.zeros_like(g0))  # grad_kappa applies kappa itself to the geometric G
    op.G = G0
    grad = op.grad_kappa(np.exp(g0), u, lam)
    h = 1e-5
    for d in range(5):
        dg = seeded_uniform(g0.size, 100 + d)
        fd = (J(g0 + h * dg) - J(g0 - h * dg)) / (2 * h)
        assert abs(fd - grad @ dg) <= 1e-6 * abs(fd), (d, fd, grad @ dg)
    assert len(free) > 0
