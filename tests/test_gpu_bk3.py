"""GPU parity of the BK3 operator (libpmg.so) against the CPU oracle.

Bar: FP64 within 1e-12 relative (max-norm) of the oracle's restatement of
BK3::Parallel::KokkosKernel + LaplaceOperatorBK3::vmult; FP32 within 2e-6.
"""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu

CASES = [
    # (cells per dir, degree, nq, deform)
    (3, 1, 2, 0), (2, 1, 3, 0), (3, 2, 3, 0), (2, 2, 4, 1), (3, 3, 4, 1), (2, 3, 5, 0),
    (4, 4, 5, 0), (3, 4, 6, 1), (2, 5, 6, 0), (2, 5, 7, 1), (2, 6, 7, 1), (2, 6, 8, 0),
]


def _pair(pmg, orc, n, p, nq, deform, precision):
    b = pmg.box(n, deform=deform, amplitude=0.05)
    m = orc.mesh(n, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
    ref = orc.LaplaceBK3(m, p, nq)
    return op, ref


@pytest.mark.parametrize("n,p,nq,deform", CASES)
def test_vmult_fp64_matches_oracle(pmg, orc, n, p, nq, deform):
    import torch
    print(f"[bk3 fp64] n={n} p={p} nq={nq} deform={deform}", flush=True)
    op, ref = _pair(pmg, orc, n, p, nq, deform, pmg.FP64)
    x = seeded_uniform(op.m(), 1000 + 10 * p + nq)
    src = torch.tensor(x, device="cuda")
    dst = op.initialize_dof_vector()
    op.vmult(dst, src)
    y = ref.vmult(x)
    assert rel_err(dst.cpu().numpy(), y) < 1e-12
    # Tvmult == vmult, and repeated application is reproducible to round-off
    dst2 = op.initialize_dof_vector()
    op.Tvmult(dst2, src)
    assert rel_err(dst2.cpu().numpy(), y) < 1e-12


@pytest.mark.parametrize("n,p,nq,deform", [(3, 1, 2, 0), (3, 2, 3, 1), (3, 4, 5, 0), (2, 6, 7, 1)])
def test_vmult_fp32_matches_oracle(pmg, orc, n, p, nq, deform):
    import torch
    print(f"[bk3 fp32] n={n} p={p} nq={nq}", flush=True)
    op, ref = _pair(pmg, orc, n, p, nq, deform, pmg.FP32)
    x = seeded_uniform(op.m(), 2000 + p)
    y = ref.vmult(x.astype(np.float32).astype(np.float64))
    # the default schedule (patch for nq <= 6 in FP32), then the atomic one
    for sched in (None, pmg.SCHEDULE_ATOMIC):
        if sched is not None:
            op.set_schedule(sched, 2)
        dst = op.initialize_dof_vector()
        op.vmult(dst, torch.tensor(x, dtype=torch.float32, device="cuda"))
        assert rel_err(dst.cpu().numpy(), y) < 2e-6, sched


def test_apply_add_and_constrained_rows(pmg, orc):
    import torch
    op, ref = _pair(pmg, orc, 3, 3, 4, 1, pmg.FP64)
    x = seeded_uniform(op.m(), 7)
    src = torch.tensor(x, device="cuda")
    base = seeded_uniform(op.m(), 8)
    dst = torch.tensor(base, device="cuda")
    op.apply_add(dst, src)
    y = ref.cells_add(x, base.copy())
    assert rel_err(dst.cpu().numpy(), y) < 1e-12
    # constrained rows of vmult are the identity
    out = op.initialize_dof_vector()
    op.vmult(out, src)
    cons = ref.cons
    assert np.array_equal(out.cpu().numpy()[cons], x[cons])


@pytest.mark.parametrize("n,p,nq,deform", [(3, 1, 2, 0), (3, 2, 3, 1), (2, 4, 5, 0), (2, 4, 6, 1),
                                           (2, 6, 7, 0)])
def test_diagonal_matches_dense(pmg, orc, n, p, nq, deform):
    op, ref = _pair(pmg, orc, n, p, nq, deform, pmg.FP64)
    op.compute_diagonal()
    inv = op.get_matrix_diagonal_inverse().cpu().numpy()
    d = ref.diagonal()
    assert rel_err(1.0 / inv, d) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vmult_above_fused_prep_threshold_matches_oracle(pmg, orc, prec):
    """A level of more than 2^20 DoFs (26^3 Q4: 1 157 625), where vmult
    zero-fills the output and copies the constrained rows in passes of their
    own: the result against the oracle, with the output vector full of NaN
    beforehand and the constrained rows equal to the source exactly."""
    import torch
    print(f"[bk3 large] 26^3 Q4 {prec}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    op, ref = _pair(pmg, orc, 26, 4, 5, 1, precision)
    assert op.m() > 2 ** 20
    x = seeded_uniform(op.m(), 2601)
    xs = x if prec == "fp64" else x.astype(np.float32).astype(np.float64)
    y = ref.vmult(xs)
    src = torch.tensor(x, dtype=dt, device="cuda")
    dst = op.initialize_dof_vector()
    dst.fill_(float("nan"))
    op.vmult(dst, src)
    out = dst.double().cpu().numpy()
    assert rel_err(out, y) < (1e-12 if prec == "fp64" else 2e-6)
    cons = pmg.box_boundary_dofs(pmg.box(26, deform=1, amplitude=0.05), 4).astype(np.int64)
    assert np.array_equal(out[cons], src.double().cpu().numpy()[cons])


def test_large_vmult_properties(pmg):
    """Size-independent checks at the 64^3 Q4 workload: symmetry <Ax,y> = <x,Ay>
    and A 1 = 0 on rows whose stencil touches no constrained DoF."""
    import torch
    b = pmg.box(64)
    op = pmg.LaplaceOperatorBK3(b, 4, 5, pmg.FP64)
    n = op.m()
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    cons = torch.tensor(pmg.box_boundary_dofs(b, 4).astype(np.int64), device="cuda")
    x[cons] = 0
    y[cons] = 0
    ax, ay = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(ax, x)
    op.vmult(ay, y)
    lhs, rhs = torch.dot(ax, y).item(), torch.dot(x, ay).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    ones = torch.ones(n, dtype=torch.float64, device="cuda")
    a1 = op.initialize_dof_vector()
    op.vmult(a1, ones)
    N = 4 * 64 + 1
    a1 = a1.reshape(N, N, N)
    inner = a1[5:-5, 5:-5, 5:-5]
    assert inner.abs().max().item() < 1e-11 * a1.abs().max().item()


def test_coefficient_and_gradient_match_oracle(pmg, orc):
    import torch
    n, p, nq = 3, 3, 4
    op, ref = _pair(pmg, orc, n, p, nq, 1, pmg.FP64)
    kappa = np.exp(seeded_uniform(op.n_cells * nq ** 3, 21, -0.5, 0.5))
    op.set_coefficient(torch.tensor(kappa, device="cuda"))
    ref.G = orc.scale_G(ref.G, kappa, nq)
    x = seeded_uniform(op.m(), 22)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(x, device="cuda"))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12
    u, lam = seeded_uniform(op.m(), 23), seeded_uniform(op.m(), 24)
    u[ref.cons] = 0
    lam[ref.cons] = 0
    out = torch.zeros(op.n_cells * nq ** 3, dtype=torch.float64, device="cuda")
    op.gradient_coefficient(torch.tensor(u, device="cuda"), torch.tensor(lam, device="cuda"), out)
    G0 = orc.geometry(ref.m, nq)[0]
    ref.G = G0
    g_ref = ref.grad_kappa(kappa, u, lam)
    assert rel_err(out.cpu().numpy(), g_ref) < 1e-12


@pytest.mark.parametrize("deform", [0, 2])
@pytest.mark.parametrize("p,nq", [(1, 2), (2, 3), (4, 5), (4, 6), (6, 7)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_affine_geometry_vmult_matches_oracle(pmg, orc, deform, p, nq, prec):
    """Affine cells (the undeformed box and the affine shear): vmult from 6
    values per cell against the oracle, which uses G at every q-point, in the
    patch and the atomic schedule."""
    import torch
    print(f"[bk3 affine] deform={deform} p={p} nq={nq} {prec}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    tol = 1e-12 if prec == "fp64" else 2e-6
    op, ref = _pair(pmg, orc, 3, p, nq, deform, precision)
    op.use_affine_geometry()
    x = seeded_uniform(op.m(), 3000 + 10 * p + nq)
    y = ref.vmult(x if prec == "fp64" else x.astype(np.float32).astype(np.float64))
    src = torch.tensor(x, dtype=dt, device="cuda")
    for sched in (pmg.SCHEDULE_PATCH, pmg.SCHEDULE_ATOMIC):
        op.set_schedule(sched, -1)
        dst = op.initialize_dof_vector()
        dst.fill_(float("nan"))
        op.vmult(dst, src)
        assert rel_err(dst.double().cpu().numpy(), y) < tol, sched


def test_affine_geometry_refused_and_turned_off(pmg, orc):
    """A deformed (non-affine) mesh is refused; a coefficient turns the
    affine form off, and vmult then uses kappa G at every q-point."""
    import torch
    print("[bk3 affine] refusal and coefficient", flush=True)
    op, _ = _pair(pmg, orc, 3, 3, 4, 1, pmg.FP64)
    with pytest.raises(pmg.PMGError) as e:
        op.use_affine_geometry()
    assert e.value.status == pmg.ERR_INVALID_ARG
    n, p, nq = 3, 3, 4
    op, ref = _pair(pmg, orc, n, p, nq, 0, pmg.FP64)
    op.use_affine_geometry()
    kappa = np.exp(seeded_uniform(op.n_cells * nq ** 3, 25, -0.5, 0.5))
    op.set_coefficient(torch.tensor(kappa, device="cuda"))
    ref.G = orc.scale_G(ref.G, kappa, nq)
    x = seeded_uniform(op.m(), 26)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(x, device="cuda"))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12
    with pytest.raises(pmg.PMGError):
        op.use_affine_geometry()  # not with a coefficient
    op.set_coefficient(None)
    op.use_affine_geometry()  # kappa = 1 again


def test_bad_arguments_raise(pmg):
    import torch
    b = pmg.box(2)
    with pytest.raises(pmg.PMGError):
        pmg.LaplaceOperatorBK3(b, 4, 9)  # unsupported nq
    op = pmg.LaplaceOperatorBK3(b, 2)
    x = torch.zeros(op.m() + 1, dtype=torch.float64, device="cuda")
    with pytest.raises(pmg.PMGError):
        op.vmult(op.initialize_dof_vector(), x)
    with pytest.raises(pmg.PMGError):
        op.get_matrix_diagonal_inverse()  # compute_diagonal() not called


@pytest.mark.parametrize("p,nq,sched,gload", [
    (4, 5, 1, 0), (4, 5, 0, 1), (4, 5, 0, 0), (4, 5, 2, 2), (4, 6, 1, 0), (4, 6, 2, 2),
    (2, 3, 1, 0), (6, 7, 1, 0), (1, 2, 2, 2), (3, 4, 2, 2), (5, 7, 2, 2), (6, 8, 2, 2),
    (1, 2, 0, 2), (3, 4, 0, 2), (4, 5, 2, 0), (4, 5, 2, 1), (3, 4, 2, 0), (6, 7, 2, 0),
    (1, 2, 2, -1), (5, 6, 0, -1), (1, 2, 3, -1), (2, 3, 3, -1), (3, 4, 3, -1), (4, 5, 3, -1),
    (4, 5, 3, 0), (4, 6, 3, -1), (5, 6, 3, -1), (6, 7, 3, -1), (6, 8, 3, -1)])
def test_schedules_agree(pmg, orc, p, nq, sched, gload):
    """Atomics-free colour-by-colour scatter, the batch-patch scatter, the
    per-cell atomic scatter, the cell-buffer gather and the G-load variants
    all give the oracle's result."""
    import torch
    op, ref = _pair(pmg, orc, 3, p, nq, 1, pmg.FP64)
    op.set_schedule(sched, gload)
    x = seeded_uniform(op.m(), 300 + p)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(x, device="cuda"))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12


@pytest.mark.parametrize("p,nq", [(4, 6), (5, 6), (5, 7), (6, 7)])
@pytest.mark.parametrize("accum", ["colored", "atomic"])
def test_patch_matrix_core_sweeps_match_oracle(pmg, orc, p, nq, accum, monkeypatch):
    """FP32 patch kernel with the x and y sweeps as batched contractions on
    the matrix cores (PMG_BK3_MFMA=1), both patch accumulations, on 27
    deformed cells (the last batch partial): the oracle's result to the FP32
    bar, and not bitwise the vector sweeps' result (the matrix path ran)."""
    import torch
    print(f"[bk3 mfma] p={p} nq={nq} accum={accum}", flush=True)
    monkeypatch.setenv("PMG_PATCH_ACCUM", accum)
    op, ref = _pair(pmg, orc, 3, p, nq, 1, pmg.FP32)
    x = seeded_uniform(op.m(), 700 + 10 * p + nq)
    y = ref.vmult(x.astype(np.float32).astype(np.float64))
    xs = torch.tensor(x, dtype=torch.float32, device="cuda")
    op.set_schedule(pmg.SCHEDULE_PATCH, 0)
    plain = op.initialize_dof_vector()
    op.vmult(plain, xs)
    monkeypatch.setenv("PMG_BK3_MFMA", "1")
    opm, _ = _pair(pmg, orc, 3, p, nq, 1, pmg.FP32)
    opm.set_schedule(pmg.SCHEDULE_PATCH, 0)
    mx = opm.initialize_dof_vector()
    for _ in range(2):  # a second application gives the same result
        opm.vmult(mx, xs)
        assert rel_err(mx.cpu().numpy(), y) < 2e-6
    assert rel_err(plain.cpu().numpy(), y) < 2e-6
    assert not torch.equal(mx, plain)


@pytest.mark.parametrize("order", ["grouped", "morton"])
def test_internal_cell_orders(pmg, orc, order, monkeypatch):
    """Internal cell reorderings (and the patch schedule on compact groups)
    leave vmult, the diagonal, the coefficient and its gradient unchanged,
    also for an operator created from a caller-given G."""
    import torch
    monkeypatch.setenv("PMG_CELL_ORDER", order)
    n, p, nq = 4, 3, 4
    op, ref = _pair(pmg, orc, n, p, nq, 1, pmg.FP64)
    x = seeded_uniform(op.m(), 61)
    y = ref.vmult(x)
    for sched in (0, 2, 3):
        op.set_schedule(sched, 2)
        dst = op.initialize_dof_vector()
        op.vmult(dst, torch.tensor(x, device="cuda"))
        assert rel_err(dst.cpu().numpy(), y) < 1e-12
    op.compute_diagonal()
    assert rel_err(1.0 / op.get_matrix_diagonal_inverse().cpu().numpy(), ref.diagonal()) < 1e-12
    kappa = np.exp(seeded_uniform(op.n_cells * nq ** 3, 62, -0.5, 0.5))
    op.set_coefficient(torch.tensor(kappa, device="cuda"))
    G0 = ref.G
    ref.G = orc.scale_G(G0, kappa, nq)
    dst = op.initialize_dof_vector()
    op.vmult(dst, torch.tensor(x, device="cuda"))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12
    u, lam = seeded_uniform(op.m(), 63), seeded_uniform(op.m(), 64)
    u[ref.cons] = 0
    lam[ref.cons] = 0
    out = torch.zeros(op.n_cells * nq ** 3, dtype=torch.float64, device="cuda")
    op.gradient_coefficient(torch.tensor(u, device="cuda"), torch.tensor(lam, device="cuda"), out)
    ref.G = G0
    assert rel_err(out.cpu().numpy(), ref.grad_kappa(kappa, u, lam)) < 1e-12
    # caller-given G (reference layout [cell][comp][q]) is re-laid out through the order
    b = op.box
    opg = pmg.LaplaceOperatorBK3.from_arrays(p, pmg.box_dofmap(b, p), op.m(),
                                             pmg.box_boundary_dofs(b, p), G=ref.G,
                                             n_q_points_1d=nq)
    dst = opg.initialize_dof_vector()
    opg.vmult(dst, torch.tensor(x, device="cuda"))
    assert rel_err(dst.cpu().numpy(), y) < 1e-12


@pytest.mark.parametrize("n,p,nq,deform", [(3, 2, 3, 1), (2, 4, 5, 0), (4, 4, 5, 1), (2, 6, 7, 1),
                                           (3, 3, 5, 2), (2, 1, 2, 1)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_bk3_rhs_and_l2_error_match_oracle(pmg, orc, n, p, nq, deform, prec):
    """assemble_rhs (f = 1 and f = 3 pi^2 sin sin sin, Gauss p+1 and nq) and
    l2_error (u = 0 and sin sin sin, Gauss p+2) against the oracle's
    quadrature sums (whose closed forms test_oracle_pinning pins): 1e-12 in
    FP64; FP32 vectors within FP32 rounding."""
    import torch
    print(f"[bk3 rhs/l2] n={n} p={p} nq={nq} deform={deform} {prec}", flush=True)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    b = pmg.box(n, deform=deform, amplitude=0.05)
    m = orc.mesh(n, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
    ref = orc.LaplaceBK3(m, p, nq)
    tol = 1e-12 if prec == "fp64" else 2e-6
    for f_kind in (0, 1):
        for q in (0, nq):
            got = op.assemble_rhs(f_kind, q).double().cpu().numpy()
            exp = ref.rhs(f_kind, None if q == 0 else q)
            assert rel_err(got, exp) < tol, (f_kind, q)
    u = seeded_uniform(op.m(), 95)
    u[orc.boundary_dofs(m, p)] = 0
    ut = torch.tensor(u, dtype=op.dtype, device="cuda")
    uq = ut.double().cpu().numpy()
    for u_kind in (0, 1):
        for q in (0, p + 3):
            e = op.l2_error(ut, u_kind, q)
            er = ref.l2_error(uq, u_kind, None if q == 0 else q)
            assert abs(e - er) <= (1e-12 if prec == "fp64" else 1e-6) * er, (u_kind, q)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_colored_schedule_is_bitwise_reproducible(pmg, prec):
    """The atomics-free colored and gather schedules add every cell's values
    in one fixed order, so repeated vmults (and a solve driven by them) are
    bitwise equal;
    the default atomic scatter may differ in the last bits between runs (its
    summation order is the hardware's), within 1e-15 relative in FP64."""
    import torch
    print(f"[determinism] {prec}", flush=True)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    b = pmg.box(6, deform=1, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, 4, None, precision)
    x = torch.tensor(seeded_uniform(op.m(), 121), dtype=op.dtype, device="cuda")
    for sched in (pmg.SCHEDULE_GATHER, pmg.SCHEDULE_COLORED):
        op.set_schedule(sched)
        ys = []
        for _ in range(4):
            y = op.initialize_dof_vector()
            op.vmult(y, x)
            ys.append(y)
        assert all(torch.equal(ys[0], y) for y in ys[1:]), sched
    op.compute_diagonal()
    sol = []
    for _ in range(2):
        u = op.initialize_dof_vector()
        pmg.solve_cg(op, x, u, pmg.PRECOND_JACOBI, None, 1e-6 if prec == "fp32" else 1e-10, 2000)
        sol.append(u)
    assert torch.equal(sol[0], sol[1])
    op.set_schedule(pmg.SCHEDULE_ATOMIC)
    ya, yb = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(ya, x)
    op.vmult(yb, x)
    tol = 1e-15 if prec == "fp64" else 1e-6
    assert rel_err(ya.double().cpu().numpy(), ys[0].double().cpu().numpy()) < tol * 10
    assert rel_err(ya.double().cpu().numpy(), yb.double().cpu().numpy()) < tol * 10


@pytest.mark.parametrize("p,nq,prec,dims", [(1, 2, "fp64", (3, 2, 4)), (2, 3, "fp64", (3, 3, 3)),
                                            (3, 4, "fp32", (4, 3, 2)), (4, 5, "fp64", (3, 4, 3)),
                                            (4, 5, "fp32", (3, 3, 3)), (4, 6, "fp64", (2, 3, 3)),
                                            (5, 6, "fp64", (2, 2, 3)), (6, 7, "fp64", (2, 2, 2)),
                                            (6, 7, "fp32", (3, 2, 2))])
def test_grid_schedule_matches_oracle(pmg, orc, p, nq, prec, dims):
    """Schedule 4 (exploded-grid store, then a sum of the <= 8 adjacent copies
    of each DoF) on a declared lexicographic box: vmult, apply_add and the
    prepared vmult of the smoother agree with the oracle, repeat bit for bit,
    and leave the constrained rows as vmult defines them."""
    import torch
    print(f"[grid schedule] p={p} nq={nq} {prec} {dims}", flush=True)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    tol = 1e-12 if prec == "fp64" else 2e-6
    b = pmg.box(*dims, deform=1, amplitude=0.05)
    m = orc.mesh(*dims, deform=1, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
    ref = orc.LaplaceBK3(m, p, nq)
    op.set_box_structure(*dims)
    op.set_schedule(pmg.SCHEDULE_GRID)
    x = seeded_uniform(op.m(), 171)
    xt = torch.tensor(x, dtype=op.dtype, device="cuda")
    y = torch.tensor(seeded_uniform(op.m(), 172), dtype=op.dtype, device="cuda")  # stale values
    op.vmult(y, xt)
    yr = ref.vmult(x)
    assert rel_err(y.double().cpu().numpy(), yr) < tol
    y2 = op.initialize_dof_vector()
    op.vmult(y2, xt)
    assert torch.equal(y, y2)
    z0 = seeded_uniform(op.m(), 173)
    z = torch.tensor(z0, dtype=op.dtype, device="cuda")
    op.apply_add(z, xt)
    cell = yr.copy()
    cons = orc.boundary_dofs(m, p)
    cell[cons] = 0.0
    assert rel_err(z.double().cpu().numpy() - z0, cell) < tol * 10
    # the smoother's prepared matvecs on this schedule
    op.compute_diagonal()
    sm = pmg.PreconditionChebyshev(op, 3, 15.0, 10)
    out_g = op.initialize_dof_vector()
    sm.vmult(out_g, xt)
    op.set_schedule(pmg.SCHEDULE_ATOMIC)
    out_a = op.initialize_dof_vector()
    sm.vmult(out_a, xt)
    assert rel_err(out_g.double().cpu().numpy(), out_a.double().cpu().numpy()) < tol * 10


def test_box_structure_is_verified(pmg):
    """set_box_structure accepts only the lexicographic box map of the given
    dimensions: other dimensions with the same counts, or a renumbered map,
    are refused and the schedule stays what it was."""
    import torch
    b = pmg.box(2, 3, 4)
    op = pmg.LaplaceOperatorBK3(b, 2)
    with pytest.raises(pmg.PMGError):
        op.set_box_structure(3, 2, 4)  # same cell count, not this numbering
    with pytest.raises(pmg.PMGError):
        op.set_box_structure(2, 3, 5)
    N, C = pmg.box_n_dofs(b, 2), pmg.box_n_cells(b)
    perm = np.random.Generator(np.random.PCG64(7)).permutation(N).astype(np.int64)
    m = pmg.box_dofmap(b, 2).astype(np.int64)
    ok = m != 0xFFFFFFFF
    m[ok] = perm[m[ok]]
    cons = perm[pmg.box_boundary_dofs(b, 2).astype(np.int64)]
    op2 = pmg.LaplaceOperatorBK3.from_arrays(2, m.astype(np.uint32), N, cons.astype(np.uint32),
                                             cell_vertices=pmg.box_cell_vertices(b))
    with pytest.raises(pmg.PMGError, match="not the lexicographic box map"):
        op2.set_box_structure(2, 3, 4)
    op.set_box_structure(2, 3, 4)  # the right one is accepted
