"""Adjoint gradient of the variable-coefficient problem on the GPU, end to end:
-div(exp(g) grad u) = 1, J = 1/2 |u - u_d|^2, dJ/dg at the Gauss points from
a forward and an adjoint MG-CG solve (FP32 V-cycle inside FP64 CG) and the
gradient kernel, wrapped in a torch.autograd.Function (test harness only).

Bars: central finite differences (step 1e-6) along random directions agree
to 1e-5 relative; on a small mesh the whole gradient matches the oracle's
(dense solves) to 1e-8 relative (limited by the 1e-12 CG tolerance)."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu


def kappa_solve():
    import torch

    class KappaPoissonSolve(torch.autograd.Function):
        """u = A(exp(g))^{-1} b; backward: lambda = A^{-1} grad_u, dJ/dg from the kernel."""

        @staticmethod
        def forward(ctx, g, prob):
            prob.set_g(g.detach())
            u = prob.solve()
            ctx.prob = prob
            ctx.save_for_backward(u)
            return u

        @staticmethod
        def backward(ctx, grad_u):
            (u,) = ctx.saved_tensors
            lam = ctx.prob.solve(grad_u.contiguous())
            return ctx.prob.gradient(u, lam), None

    return KappaPoissonSolve


def _problem(pmg, n, p, n_levels, seed=5, tol=1e-12):
    import torch
    g_smooth = pmg.random_fourier_field(seed)
    prob = pmg.CoefficientPoisson(pmg.box(n), p, n_levels, g_smooth, pmg.FP32, rel_tol=tol)
    prob.set_g(torch.zeros(len(prob.X), dtype=torch.float64, device="cuda"))
    u_d = prob.solve()
    g0 = torch.tensor(g_smooth(prob.X), device="cuda")
    return prob, u_d, g0


def test_adjoint_gradient_matches_central_differences(pmg):
    import torch
    print("[adjoint] 8^3 Q4, FP32 MG in FP64 CG", flush=True)
    prob, u_d, g0 = _problem(pmg, 8, 4, 3)
    Solve = kappa_solve()
    g = g0.clone().requires_grad_(True)
    u = Solve.apply(g, prob)
    J = 0.5 * torch.sum((u - u_d) ** 2)
    J.backward()
    grad = g.grad.detach()

    def J_at(gv):
        prob.set_g(gv)
        return 0.5 * float(torch.sum((prob.solve() - u_d) ** 2))

    h = 1e-6
    for d in range(5):
        dg = torch.tensor(seeded_uniform(g0.numel(), 200 + d), device="cuda")
        fd = (J_at(g0 + h * dg) - J_at(g0 - h * dg)) / (2 * h)
        an = float(torch.dot(grad, dg))
        assert abs(fd - an) <= 1e-5 * abs(fd), (d, fd, an)


def test_adjoint_gradient_matches_oracle(pmg, orc):
    """4^3 Q2: the autograd gradient against the oracle's dense-solve gradient."""
    import torch
    n, p, nq = 4, 2, 3
    prob, u_d, g0 = _problem(pmg, n, p, 2)
    Solve = kappa_solve()
    g = g0.clone().requires_grad_(True)
    u = Solve.apply(g, prob)
    (0.5 * torch.sum((u - u_d) ** 2)).backward()

    m = orc.mesh(n)
    G0 = orc.geometry(m, nq)[0]
    base = orc.LaplaceBK3(m, p, nq)

    def solve(gv, rhs):
        op = orc.LaplaceBK3(m, p, nq, G=orc.scale_G(G0, np.exp(gv), nq))
        A = np.stack([op.vmult(e) for e in np.eye(base.N)], axis=1)
        return np.linalg.solve(A, rhs)

    b = base.rhs(0)
    gn = g0.cpu().numpy()
    ud = solve(np.zeros_like(gn), b)
    uo = solve(gn, b)
    lam = solve(gn, uo - ud)
    grad_ref = base.grad_kappa(np.exp(gn), uo, lam)
    assert rel_err(u.detach().cpu().numpy(), uo) < 1e-10
    assert rel_err(g.grad.cpu().numpy(), grad_ref) < 1e-8
