"""GPU parity of transfers, Chebyshev smoother, V-cycle and (mixed-precision)
MG-CG against the CPU oracle.

Bars: transfers and smoother applications within 1e-12 relative (FP64) and
2e-6 (FP32); V-cycle within 1e-9 relative (its coarse solve is an iterative
1e-12 PCG on both sides); CG iteration counts within +-1 of the oracle and
solutions within 1e-9 relative; the FP32 V-cycle inside FP64 CG reaches the
same 1e-10 residual and a solution within 1e-8 of the all-FP64 one.
"""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


@pytest.mark.parametrize("n,p,deform", [(2, 1, 0), (2, 2, 1), (2, 4, 0), (1, 6, 0), (3, 3, 1)])
def test_h_transfer_matches_oracle(pmg, orc, n, p, deform):
    import torch
    print(f"[h-transfer] n={n} p={p}", flush=True)
    bc = pmg.box(n, deform=deform, amplitude=0.05)
    tr = pmg.GeometricTransfer(bc, p)
    ref = orc.Transfer(orc.mesh(n, deform=deform, amplitude=0.05), p, kind="h")
    nc, nf = pmg.box_n_dofs(bc, p), pmg.box_n_dofs(pmg.refine(bc), p)
    xc, xf = seeded_uniform(nc, 31), seeded_uniform(nf, 32)
    yf0, yc0 = seeded_uniform(nf, 33), seeded_uniform(nc, 34)
    yf = _t(yf0)
    tr.prolongate_and_add(yf, _t(xc))
    assert rel_err(yf.cpu().numpy(), ref.prolongate_add(xc, yf0.copy())) < 1e-12
    yc = _t(yc0)
    tr.restrict_and_add(yc, _t(xf))
    assert rel_err(yc.cpu().numpy(), ref.restrict_add(xf, yc0.copy())) < 1e-12


@pytest.mark.parametrize("pc,pf", [(1, 2), (2, 4), (1, 3), (3, 6), (4, 6), (5, 6), (2, 3)])
def test_p_transfer_matches_oracle(pmg, orc, pc, pf):
    print(f"[p-transfer] {pc}->{pf}", flush=True)
    b = pmg.box(3, deform=1, amplitude=0.05)
    tr = pmg.PolynomialTransfer(b, pc, pf)
    ref = orc.Transfer(orc.mesh(3, deform=1, amplitude=0.05), pc, pf, kind="p")
    nc, nf = pmg.box_n_dofs(b, pc), pmg.box_n_dofs(b, pf)
    xc, xf = seeded_uniform(nc, 41), seeded_uniform(nf, 42)
    yf = _t(np.zeros(nf))
    tr.prolongate_and_add(yf, _t(xc))
    assert rel_err(yf.cpu().numpy(), ref.prolongate_add(xc, np.zeros(nf))) < 1e-12
    yc = _t(np.zeros(nc))
    tr.restrict_and_add(yc, _t(xf))
    assert rel_err(yc.cpu().numpy(), ref.restrict_add(xf, np.zeros(nc))) < 1e-12


@pytest.mark.parametrize("gather", ["0", "1"])
@pytest.mark.parametrize("kind,prec", [("h", "fp64"), ("h", "fp32"), ("p", "fp64"), ("p", "fp32")])
def test_restriction_paths_match_oracle(pmg, orc, monkeypatch, gather, kind, prec):
    """Both restriction paths: coarse FP atomics, and cell values stored to a
    buffer then gathered through the inverse coarse map (used on large levels)."""
    import torch
    monkeypatch.setenv("PMG_RESTRICT_GATHER", gather)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    tol = 1e-12 if prec == "fp64" else 2e-6
    b = pmg.box(3, deform=1, amplitude=0.05)
    if kind == "h":
        tr = pmg.GeometricTransfer(b, 2, precision=precision)
        ref = orc.Transfer(orc.mesh(3, deform=1, amplitude=0.05), 2, kind="h")
        nc, nf = pmg.box_n_dofs(b, 2), pmg.box_n_dofs(pmg.refine(b), 2)
    else:
        tr = pmg.PolynomialTransfer(b, 2, 4, precision=precision)
        ref = orc.Transfer(orc.mesh(3, deform=1, amplitude=0.05), 2, 4, kind="p")
        nc, nf = pmg.box_n_dofs(b, 2), pmg.box_n_dofs(b, 4)
    xf, yc0 = seeded_uniform(nf, 45), seeded_uniform(nc, 46)
    for rep in range(2):  # the buffer is reused by the second call
        yc = _t(yc0, dt)
        tr.restrict_and_add(yc, _t(xf, dt))
        assert rel_err(yc.double().cpu().numpy(), ref.restrict_add(xf, yc0.copy())) < tol


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_restriction_buffer_sums_long_ranges(pmg, monkeypatch, prec):
    """A coarse map in which one DoF collects the entries of six interior
    vertices (48 entries: more than the 8 the buffer sum loads without
    branches) and five DoFs have none: the buffered restriction equals the
    atomic one, on two calls in a row."""
    import torch
    print(f"[restriction buffer] long ranges {prec}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    tol = 1e-12 if prec == "fp64" else 2e-6
    b = pmg.box(4, deform=1, amplitude=0.05)
    pc, pf = 1, 3
    mc, mf, _ = pmg.box_p_transfer_maps(b, pc, pf, weights=False)
    n_c, n_f = pmg.box_n_dofs(b, pc), pmg.box_n_dofs(b, pf)
    cons = pmg.box_boundary_dofs(b, pf)
    valid = mc != np.uint32(0xFFFFFFFF)
    counts = np.bincount(mc[valid].astype(np.int64), minlength=n_c)
    glue = np.flatnonzero(counts == 8)[:6]
    assert glue.size == 6
    mc = mc.copy()
    mc[np.isin(mc, glue)] = glue[0]
    assert np.count_nonzero(mc == glue[0]) == 48
    monkeypatch.setenv("PMG_RESTRICT_GATHER", "0")
    ta = pmg._Transfer(pmg.TRANSFER_P, precision, pc, pf, mc, mf, None, n_c, n_f, 1.0, cons=cons)
    monkeypatch.setenv("PMG_RESTRICT_GATHER", "1")
    tg = pmg._Transfer(pmg.TRANSFER_P, precision, pc, pf, mc, mf, None, n_c, n_f, 1.0, cons=cons)
    assert ta.index_data("restrict_offsets").size == 0
    off = tg.index_data("restrict_offsets")
    assert int(off[glue[0] + 1]) - int(off[glue[0]]) == 48
    for seed in (47, 48):
        xf, yc0 = seeded_uniform(n_f, seed), seeded_uniform(n_c, seed + 10)
        ya, yg = _t(yc0, dt), _t(yc0, dt)
        ta.restrict_and_add(ya, _t(xf, dt))
        tg.restrict_and_add(yg, _t(xf, dt))
        a, g = ya.double().cpu().numpy(), yg.double().cpu().numpy()
        assert rel_err(g, a) < tol
        y0 = _t(yc0, dt).double().cpu().numpy()
        assert np.array_equal(g[glue[1:]], y0[glue[1:]])  # DoFs without entries


def test_transfer_fp32_and_doubled_prolongation(pmg, orc):
    import torch
    bc = pmg.box(2)
    ref = orc.Transfer(orc.mesh(2), 2, kind="h", h_prolong_scale=2.0)
    nc, nf = pmg.box_n_dofs(bc, 2), pmg.box_n_dofs(pmg.refine(bc), 2)
    xc = seeded_uniform(nc, 51)
    tr2 = pmg.GeometricTransfer(bc, 2, prolongation_scale=2.0)
    yf = _t(np.zeros(nf))
    tr2.prolongate_and_add(yf, _t(xc))
    assert rel_err(yf.cpu().numpy(), ref.prolongate_add(xc, np.zeros(nf))) < 1e-12
    tr32 = pmg.GeometricTransfer(bc, 2, precision=pmg.FP32)
    y32 = torch.zeros(nf, dtype=torch.float32, device="cuda")
    tr32.prolongate_and_add(y32, _t(xc, torch.float32))
    assert rel_err(2.0 * y32.double().cpu().numpy(), ref.prolongate_add(xc, np.zeros(nf))) < 2e-6


def _oracle_levels(orc, boxes_pm):
    return [orc.mesh(b.nx, b.ny, b.nz, deform=b.deform, amplitude=b.amplitude) for b in boxes_pm]


@pytest.mark.parametrize("p,deform", [(1, 0), (2, 1), (4, 0)])
def test_chebyshev_matches_oracle(pmg, orc, p, deform):
    print(f"[chebyshev] p={p}", flush=True)
    b = pmg.box(4, deform=deform, amplitude=0.05)
    m = orc.mesh(4, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p)
    op.compute_diagonal()
    sm = pmg.PreconditionChebyshev(op, 3, 15.0, 10)
    ref = orc.Multigrid([orc.mesh(2, deform=deform, amplitude=0.05), m], [p, p])
    mx, mn, deg = sm.info()
    rmx, rmn, rdeg = ref.smoother_info(1)
    assert deg == rdeg == 3
    assert abs(mx - rmx) <= 1e-10 * rmx and abs(mn - rmn) <= 1e-8 * abs(rmn)
    rhs = seeded_uniform(op.m(), 61)
    x = op.initialize_dof_vector()
    sm.vmult(x, _t(rhs))
    xr = ref.smooth(1, np.zeros(op.m()), rhs, 0)
    assert rel_err(x.cpu().numpy(), xr) < 1e-11
    sm.step(x, _t(rhs))
    xr = ref.smooth(1, xr, rhs, 1)
    assert rel_err(x.cpu().numpy(), xr) < 1e-11


@pytest.mark.parametrize("sched", ["patch", "atomic", "gather"])
@pytest.mark.parametrize("constrained", [True, False])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_chebyshev_fused_output_prep(pmg, constrained, prec, sched):
    """The smoother's updates write the next matvec's output (src on
    constrained rows, 0 elsewhere) and then run only the cell kernel.  Compare
    with the first-kind recurrence driven from the host through the operator's
    own vmult (which prepares its output itself), with and without constrained
    rows (mask present or absent)."""
    import torch
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    tol = 1e-12 if prec == "fp64" else 2e-5
    b, p = pmg.box(3, deform=1, amplitude=0.05), 2
    if constrained:
        op = pmg.LaplaceOperatorBK3(b, p, precision=precision)
    else:
        op = pmg.LaplaceOperatorBK3.from_arrays(p, pmg.box_dofmap(b, p, False), pmg.box_n_dofs(b, p),
                                                np.zeros(0, dtype=np.uint32),
                                                cell_vertices=pmg.box_cell_vertices(b),
                                                precision=precision)
    op.set_schedule({"patch": pmg.SCHEDULE_PATCH, "atomic": pmg.SCHEDULE_ATOMIC,
                     "gather": pmg.SCHEDULE_GATHER}[sched])
    op.compute_diagonal()
    sm = pmg.PreconditionChebyshev(op, 3, 15.0, 10)
    mx, _, deg = sm.info()
    alpha = mx / 15.0
    theta, delta = (mx + alpha) / 2, (mx - alpha) / 2
    P = op.get_matrix_diagonal_inverse().double()
    rhs = torch.tensor(seeded_uniform(op.m(), 62), dtype=torch.float64, device="cuda")

    def A(v):
        out = op.initialize_dof_vector()
        op.vmult(out, v.to(dt))
        return out.double()

    def host(x0, step):
        x = x0.clone()
        x = x + P * (rhs - A(x)) / theta if step else P * rhs / theta
        old, rho_k, sigma = x0.clone(), delta / theta, theta / delta
        for _ in range(deg - 1):
            rho_n = 1.0 / (2.0 * sigma - rho_k)
            x, old = (1 + rho_n * rho_k) * x - rho_n * rho_k * old + 2 * rho_n / delta * P * (rhs - A(x)), x
            rho_k = rho_n
        return x

    x = op.initialize_dof_vector()
    sm.vmult(x, rhs.to(dt))
    xh = host(torch.zeros_like(rhs), False)
    assert rel_err(x.double().cpu().numpy(), xh.cpu().numpy()) < tol
    x0 = x.double().clone()
    sm.step(x, rhs.to(dt))
    xh = host(x0, True)
    assert rel_err(x.double().cpu().numpy(), xh.cpu().numpy()) < tol


def test_coarse_chebyshev_auto_degree(pmg, orc):
    """Coarse smoother with automatic degree (range 1e-3, eig its = m())."""
    b = pmg.box(2)
    op = pmg.LaplaceOperatorBK3(b, 2)
    op.compute_diagonal()
    sm = pmg.PreconditionChebyshev(op, 0, 1e-3, op.m())
    ref = orc.Multigrid([orc.mesh(2)], [2], coarse_kind=0)
    mx, mn, deg = sm.info()
    rmx, rmn, rdeg = ref.smoother_info(0)
    assert deg == rdeg and deg > 3
    assert abs(mx - rmx) <= 1e-8 * rmx and abs(mn - rmn) <= 1e-6 * rmn


@pytest.mark.parametrize("levels_spec", [
    [(2, 2), (4, 2), (8, 2)],                 # h-only, Q2
    [(2, 1), (4, 1), (4, 2), (4, 4)],         # ph: h at Q1 then p 1->2->4
    [(2, 3), (4, 3)],                         # two-level Q3
])
def test_vcycle_and_mgcg_match_oracle(pmg, orc, levels_spec):
    import torch
    print(f"[mg] {levels_spec}", flush=True)
    boxes = [pmg.box(n, deform=1, amplitude=0.05) for n, _ in levels_spec]
    degs = [p for _, p in levels_spec]
    mg, ops, trs, sms = pmg.build_hierarchy(list(zip(boxes, degs)))
    ref = orc.Multigrid([orc.mesh(n, deform=1, amplitude=0.05) for n, _ in levels_spec], degs,
                        coarse_kind=1, coarse_tol=1e-12)
    fine = ops[-1]
    r = seeded_uniform(fine.m(), 71)
    r[orc.boundary_dofs(orc.mesh(levels_spec[-1][0]), degs[-1])] = 0
    z = fine.initialize_dof_vector()
    mg.vmult(z, _t(r))
    zr = ref.vmult(r)
    assert rel_err(z.cpu().numpy(), zr) < 1e-9
    # MG-preconditioned CG
    b = fine.assemble_rhs(1)
    x = fine.initialize_dof_vector()
    its, r0, res = pmg.solve_cg(fine, b, x, pmg.PRECOND_MG, mg, 1e-10, 200)
    xr, rits, rr0, rres = ref.pcg(b.cpu().numpy(), precond=2, rel_tol=1e-10)
    assert abs(its - rits) <= 1, (its, rits)
    assert res <= 1e-10 * r0
    assert rel_err(x.cpu().numpy(), xr) < 1e-9
    # mixed precision: FP32 V-cycle inside FP64 CG
    mg32, ops32, _, _ = pmg.build_hierarchy(list(zip(boxes, degs)), precision=pmg.FP32,
                                            outer_precision=pmg.FP64)
    x32 = fine.initialize_dof_vector()
    its32, _, res32 = pmg.solve_cg(fine, b, x32, pmg.PRECOND_MG, mg32, 1e-10, 200)
    assert res32 <= 1e-10 * r0 and its32 <= its + 2
    assert rel_err(x32.cpu().numpy(), x.cpu().numpy()) < 1e-8


def test_mixed_vcycle_keeps_coarse_solve_and_diagonal_in_fp64(pmg, orc):
    """FP32 levels under an FP64 outer solve: the coarsest operator is FP64
    (the coarse-grid solve runs in FP64), the FP32 levels' inverse diagonals
    are assembled and inverted in FP64 and rounded once, and the V-cycle
    agrees with the FP64 oracle to FP32 accuracy; an FP32 coarse level asked
    for explicitly still works."""
    import torch
    spec = [(2, 2), (4, 2), (8, 2)]
    print(f"[mg mixed coarse] {spec}", flush=True)
    boxes = [pmg.box(n, deform=1, amplitude=0.05) for n, _ in spec]
    levels = list(zip(boxes, [p for _, p in spec]))
    mg, ops, _, _ = pmg.build_hierarchy(levels, precision=pmg.FP32, outer_precision=pmg.FP64)
    assert ops[0].precision == pmg.FP64 and all(o.precision == pmg.FP32 for o in ops[1:])
    op64 = pmg.LaplaceOperatorBK3(boxes[-1], 2, None, pmg.FP64)
    op64.compute_diagonal()
    inv64 = op64.get_matrix_diagonal_inverse().cpu().numpy()
    inv32 = ops[-1].get_matrix_diagonal_inverse().cpu().numpy().astype(np.float64)
    assert rel_err(inv32, inv64) < 3e-7
    ref = orc.Multigrid([orc.mesh(n, deform=1, amplitude=0.05) for n, _ in spec],
                        [p for _, p in spec], coarse_kind=1, coarse_tol=1e-12)
    r = seeded_uniform(ops[-1].m(), 73)
    r[orc.boundary_dofs(orc.mesh(spec[-1][0]), spec[-1][1])] = 0
    zr = ref.vmult(r)
    z = torch.zeros(len(r), dtype=torch.float64, device="cuda")
    mg.vmult(z, _t(r))
    assert rel_err(z.cpu().numpy(), zr) < 1e-4
    mg_c, ops_c, _, _ = pmg.build_hierarchy(levels, precision=pmg.FP32, outer_precision=pmg.FP64,
                                            coarse_precision=pmg.FP32)
    assert ops_c[0].precision == pmg.FP32
    z.zero_()
    mg_c.vmult(z, _t(r))
    assert rel_err(z.cpu().numpy(), zr) < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vcycle_with_gathered_restriction_matches_oracle(pmg, orc, monkeypatch, prec):
    """The V-cycle with every restriction through the cell buffer and the
    inverse-map gather, forming the defect d - A s inside the restriction."""
    import torch
    monkeypatch.setenv("PMG_RESTRICT_GATHER", "1")
    spec = [(2, 2), (4, 2), (4, 3)]
    boxes = [pmg.box(n, deform=1, amplitude=0.05) for n, _ in spec]
    degs = [p for _, p in spec]
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    mg, ops, _, _ = pmg.build_hierarchy(list(zip(boxes, degs)), precision=precision,
                                        outer_precision=pmg.FP64)
    ref = orc.Multigrid([orc.mesh(n, deform=1, amplitude=0.05) for n, _ in spec], degs,
                        coarse_kind=1, coarse_tol=1e-12)
    r = seeded_uniform(pmg.box_n_dofs(boxes[-1], degs[-1]), 72)
    r[orc.boundary_dofs(orc.mesh(spec[-1][0]), degs[-1])] = 0
    z = torch.zeros(len(r), dtype=torch.float64, device="cuda")
    mg.vmult(z, _t(r))
    zr = ref.vmult(r)
    assert rel_err(z.cpu().numpy(), zr) < (1e-9 if prec == "fp64" else 1e-4)


def test_jacobi_pcg_matches_oracle(pmg, orc):
    """Small version of the Jacobi-PCG configuration: seeded U[-1,1] RHS."""
    b = pmg.box(4)
    op = pmg.LaplaceOperatorBK3(b, 4, 6)
    op.compute_diagonal()
    ref = orc.Multigrid([orc.mesh(4)], [4], nqs=[6])
    rhs = seeded_uniform(op.m(), 81)
    rhs[orc.boundary_dofs(orc.mesh(4), 4)] = 0
    x = op.initialize_dof_vector()
    its, r0, res = pmg.solve_cg(op, _t(rhs), x, pmg.PRECOND_JACOBI, None, 1e-8, 2000)
    xr, rits, _, _ = ref.pcg(rhs, level=0, precond=1, rel_tol=1e-8, max_it=2000)
    assert abs(its - rits) <= 1, (its, rits)
    # both iterates stop at 1e-8 relative residual; they agree far below that
    assert res <= 1e-8 * r0
    assert rel_err(x.cpu().numpy(), xr) < 1e-7


@pytest.mark.parametrize("every", ["1", "5", "8"])
def test_cg_result_does_not_depend_on_the_check_interval(pmg, monkeypatch, every):
    """Iterations queued past the converged one change nothing: the iteration
    count and the solution are those of a check after every iteration."""
    print(f"[cg] check every {every}", flush=True)
    b = pmg.box(4, deform=1, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, 3, 4)
    op.compute_diagonal()
    rhs = _t(seeded_uniform(op.m(), 83))
    monkeypatch.setenv("PMG_CG_CHECK_EVERY", "1")
    x1 = op.initialize_dof_vector()
    its1, r01, res1 = pmg.solve_cg(op, rhs, x1, pmg.PRECOND_JACOBI, None, 1e-9, 2000)
    monkeypatch.setenv("PMG_CG_CHECK_EVERY", every)
    x2 = op.initialize_dof_vector()
    its2, r02, res2 = pmg.solve_cg(op, rhs, x2, pmg.PRECOND_JACOBI, None, 1e-9, 2000)
    # equal up to the round-off of the FP atomics, which differs run to run;
    # one more iteration would move x by about the 1e-9 tolerance
    assert its2 == its1
    assert abs(res2 - res1) <= 1e-12 * r01
    assert rel_err(x2.cpu().numpy(), x1.cpu().numpy()) < 1e-11
    # not converged within max_it: the count is max_it either way
    with pytest.raises(pmg.PMGError) as e:
        pmg.solve_cg(op, rhs, op.initialize_dof_vector(), pmg.PRECOND_JACOBI, None, 1e-9, its1 - 3)
    assert e.value.status == pmg.ERR_NOT_CONVERGED


def test_coarse_chebyshev_vcycle_runs(pmg):
    """The reference's coarse solver (Chebyshev, automatic degree to 1e-3)."""
    boxes = [pmg.box(2), pmg.box(4)]
    mg, ops, _, _ = pmg.build_hierarchy([(boxes[0], 2), (boxes[1], 2)], coarse=pmg.COARSE_CHEBYSHEV)
    b = ops[-1].assemble_rhs(0)
    x = ops[-1].initialize_dof_vector()
    its, r0, res = pmg.solve_cg(ops[-1], b, x, pmg.PRECOND_MG, mg, 1e-10, 100)
    assert res <= 1e-10 * r0 and its < 30


def test_cg_reports_non_convergence(pmg):
    b = pmg.box(4)
    op = pmg.LaplaceOperatorBK3(b, 2)
    op.compute_diagonal()
    rhs = op.assemble_rhs(1)
    x = op.initialize_dof_vector()
    with pytest.raises(pmg.PMGError) as e:
        pmg.solve_cg(op, rhs, x, pmg.PRECOND_JACOBI, None, 1e-14, 3)
    assert e.value.status == pmg.ERR_NOT_CONVERGED


def test_cg_rejects_mismatched_mg_and_reuses_its_vectors(pmg):
    """An MG preconditioner whose finest level is not the operator's size is
    an error; repeated solves on one operator (whose handle keeps the CG
    vectors) give the same result, also after a solve of another size (to
    rounding: the atomic scatter's summation order varies between runs)."""
    import torch
    mg, ops, _, _ = pmg.build_hierarchy([(pmg.box(2), 2), (pmg.box(4), 2)])
    other = pmg.LaplaceOperatorBK3(pmg.box(3), 2)
    x = other.initialize_dof_vector()
    with pytest.raises(pmg.PMGError) as e:
        pmg.solve_cg(other, other.assemble_rhs(1), x, pmg.PRECOND_MG, mg, 1e-10, 50)
    assert e.value.status == pmg.ERR_SIZE
    fine = ops[-1]
    b = fine.assemble_rhs(1)
    runs = []
    for _ in range(2):
        x = fine.initialize_dof_vector()
        runs.append((pmg.solve_cg(fine, b, x, pmg.PRECOND_MG, mg, 1e-10, 50), x))
        other.compute_diagonal()
        y = other.initialize_dof_vector()
        pmg.solve_cg(other, other.assemble_rhs(1), y, pmg.PRECOND_JACOBI, None, 1e-8, 500)
    (it0, r00, r0), x0 = runs[0]
    (it1, r01, r1), x1 = runs[1]
    assert it0 == it1 and abs(r00 - r01) <= 1e-12 * r00 and r1 <= 1e-10 * r01
    assert rel_err(x1.cpu().numpy(), x0.cpu().numpy()) < 1e-12


def test_vcycle_is_finite_over_poisoned_memory(pmg):
    """Work vectors come from fresh device allocations whose contents are
    arbitrary: the V-cycle (including the one-workgroup coarse PCG) must never
    read them before writing.  Fill freed device memory with NaN first."""
    import torch
    junk = torch.full((1 << 27,), float("nan"), dtype=torch.float64, device="cuda")
    del junk
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    for prec in (pmg.FP64, pmg.FP32):
        mg, ops, _, _, _ = pmg.geometric_hierarchy(pmg.box(16), 4, 4, precision=prec,
                                                   outer_precision=pmg.FP64)
        r = torch.rand(ops[-1].m(), dtype=torch.float64, device="cuda")
        z = torch.zeros_like(r)
        for _ in range(3):
            mg.vmult(z, r)
            assert torch.isfinite(z).all().item()


@pytest.mark.parametrize("kind", ["h", "p"])
def test_weighted_transfer_path(pmg, orc, kind):
    """Weights other than 1/multiplicity take the weighted (atomic) kernels;
    1/multiplicity weights take the owner-computes kernels.  Both against the
    oracle's weighted transfer with the same weights."""
    m = orc.mesh(3, deform=1, amplitude=0.05)
    ref = orc.Transfer(m, 2, 4 if kind == "p" else None, kind=kind)
    b = pmg.box(3, deform=1, amplitude=0.05)
    nc = pmg.box_n_dofs(b, 2)
    nf = pmg.box_n_dofs(pmg.refine(b) if kind == "h" else b, 2 if kind == "h" else 4)
    xc, xf = seeded_uniform(nc, 301), seeded_uniform(nf, 302)
    for wscale in (1.0, 0.7):
        ref.w = ref.w * wscale
        tr = pmg._Transfer(pmg.TRANSFER_H if kind == "h" else pmg.TRANSFER_P, pmg.FP64, 2,
                           2 if kind == "h" else 4, ref.map_c, ref.map_f, ref.w, nc, nf, 1.0)
        yf = _t(np.zeros(nf))
        tr.prolongate_and_add(yf, _t(xc))
        assert rel_err(yf.cpu().numpy(), ref.prolongate_add(xc, np.zeros(nf))) < 1e-12
        yc = _t(np.zeros(nc))
        tr.restrict_and_add(yc, _t(xf))
        assert rel_err(yc.cpu().numpy(), ref.restrict_add(xf, np.zeros(nc))) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vcycle_graph_replays_match_eager(pmg, monkeypatch, prec):
    """The captured V-cycle graph, replayed several times with eager work and
    synchronisations in between, gives the eager V-cycle's result each time."""
    import torch
    p = pmg.FP64 if prec == "fp64" else pmg.FP32
    levels = pmg.ph_levels(pmg.box(8, deform=1, amplitude=0.05), 4, 3)
    mg_g, ops_g, _, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    monkeypatch.setenv("PMG_VCYCLE_GRAPH", "0")
    mg_e, ops_e, _, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    assert mg_g.uses_graph() and not mg_e.uses_graph()
    n = ops_g[-1].m()
    tol = 1e-12 if prec == "fp64" else 1e-5
    for rep in range(4):
        r = _t(seeded_uniform(n, 400 + rep))
        zg, ze = torch.zeros_like(r), torch.zeros_like(r)
        mg_g.vmult(zg, r)
        junk = torch.randn(1 << 20, dtype=torch.float64, device="cuda").sum()  # eager work
        torch.cuda.synchronize()
        mg_e.vmult(ze, r)
        torch.cuda.synchronize()
        assert rel_err(zg.cpu().numpy(), ze.cpu().numpy()) < tol, rep
        del junk


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vcycle_graph_sees_level_data_changed_after_capture(pmg, monkeypatch, prec):
    """Level data changed behind the same handles after the graph was captured
    (a new coefficient and a recomputed inverse diagonal) is what the replay
    uses: the graph keeps pointing at the buffers the operator refills."""
    import torch
    print(f"[graph after set_coefficient] {prec}", flush=True)
    p = pmg.FP64 if prec == "fp64" else pmg.FP32
    levels = pmg.ph_levels(pmg.box(8, deform=1, amplitude=0.05), 4, 3)
    mg_g, ops_g, _, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    monkeypatch.setenv("PMG_VCYCLE_GRAPH", "0")
    mg_e, ops_e, _, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    assert mg_g.uses_graph() and not mg_e.uses_graph()
    n = ops_g[-1].m()
    r = _t(seeded_uniform(n, 410))
    z0 = torch.zeros_like(r)
    mg_g.vmult(z0, r)  # captures the graph
    torch.cuda.synchronize()
    for ops in (ops_g, ops_e):
        for lev in (ops[-1], ops[-2]):
            kappa = torch.full((lev.n_cells * lev.nq ** 3,), 2.5, dtype=torch.float64,
                               device="cuda")
            lev.set_coefficient(kappa)
            lev.compute_diagonal()
    tol = 1e-12 if prec == "fp64" else 1e-5
    for rep in range(2):
        zg, ze = torch.zeros_like(r), torch.zeros_like(r)
        mg_g.vmult(zg, r)
        torch.cuda.synchronize()
        mg_e.vmult(ze, r)
        torch.cuda.synchronize()
        assert rel_err(zg.cpu().numpy(), ze.cpu().numpy()) < tol, rep
    # the changed levels do change the V-cycle (kappa 2.5 on the two finest)
    assert rel_err(zg.cpu().numpy(), z0.cpu().numpy()) > 1e-3


@pytest.mark.parametrize("kind,p,pf", [("h", 1, 1), ("h", 2, 2), ("h", 4, 4), ("p", 1, 2),
                                       ("p", 2, 5), ("p", 3, 6)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_transfer_weights_built_on_device_match_host_weights(pmg, kind, p, pf, prec):
    """setup_weights on the device (from the fine constraints) gives the
    transfer that the host weights give, whose parity with the oracle the
    other transfer tests pin."""
    import torch
    print(f"[device weights] {kind} p={p} pf={pf} {prec}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    b = pmg.box(3, deform=1, amplitude=0.05)
    if kind == "h":
        a = pmg.GeometricTransfer(b, p, precision)
        d = pmg.GeometricTransfer(b, p, precision, device_weights=True)
    else:
        a = pmg.PolynomialTransfer(b, p, pf, precision)
        d = pmg.PolynomialTransfer(b, p, pf, precision, device_weights=True)
    xc = torch.tensor(seeded_uniform(a.n_c, 81), dtype=dt, device="cuda")
    xf = torch.tensor(seeded_uniform(a.n_f, 82), dtype=dt, device="cuda")
    ya, yd = (torch.zeros(a.n_f, dtype=dt, device="cuda") for _ in range(2))
    a.prolongate_and_add(ya, xc)
    d.prolongate_and_add(yd, xc)
    tol = 1e-14 if prec == "fp64" else 1e-6
    assert rel_err(yd.cpu().numpy(), ya.cpu().numpy()) < tol
    ra, rd = (torch.zeros(a.n_c, dtype=dt, device="cuda") for _ in range(2))
    a.restrict_and_add(ra, xf)
    d.restrict_and_add(rd, xf)
    assert rel_err(rd.cpu().numpy(), ra.cpu().numpy()) < tol


def test_transfer_from_constraints_errors(pmg):
    b = pmg.box(2)
    mc, mf, _ = pmg.box_h_transfer_maps(b, 2)
    n_c, n_f = pmg.box_n_dofs(b, 2), pmg.box_n_dofs(pmg.refine(b), 2)
    with pytest.raises(pmg.PMGError, match="out of range"):
        pmg._Transfer(pmg.TRANSFER_H, pmg.FP64, 2, 2, mc, mf, None, n_c, n_f, 1.0,
                      cons=np.array([n_f], np.uint32))
    with pytest.raises(pmg.PMGError, match="h- and p-transfers"):
        pmg._Transfer(pmg.TRANSFER_DG_CG, pmg.FP64, 2, 2, pmg.box_dofmap(b, 2), None, None, n_c,
                      pmg.box_n_cells(b) * 27, 1.0, cons=np.zeros(0, np.uint32))


def test_config2_q4_h_hierarchy_matches_oracle(pmg, orc):
    """Geometric MG-CG configured as the 64^3 Q4 workload, at 8^3: Q4 on every
    h-level (2^3, 4^3, 8^3), Gauss nq = 5, Chebyshev degree 3 over the Jacobi
    diagonal (range 15, 10 eigenvalue steps), coarse Jacobi-PCG to 1e-12,
    f = 3 pi^2 sin sin sin, CG to 1e-10: the V-cycle, the iteration count
    (+-1), the solution and the L2 error agree with the oracle; the FP32
    V-cycle inside FP64 CG reaches the same residual and error."""
    import torch
    print("[config2 8^3 Q4]", flush=True)
    mg, ops, _, _, boxes = pmg.geometric_hierarchy(pmg.box(8), 4, 3, 5)
    ref = orc.Multigrid([orc.mesh(2), orc.mesh(4), orc.mesh(8)], [4, 4, 4], nqs=[5, 5, 5],
                        cheby_degree=3, smoothing_range=15.0, eig_cg_n_iterations=10,
                        coarse_kind=1, coarse_tol=1e-12)
    fine = ops[-1]
    for l in range(1, 3):  # the smoothers' eigenvalue estimates
        mx, mn, deg = pmg.PreconditionChebyshev(ops[l], 3, 15.0, 10).info()
        rmx, rmn, rdeg = ref.smoother_info(l)
        assert deg == rdeg == 3 and abs(mx - rmx) <= 1e-10 * rmx and abs(mn - rmn) <= 1e-9 * rmx
    r = seeded_uniform(fine.m(), 72)
    r[orc.boundary_dofs(orc.mesh(8), 4)] = 0
    z = fine.initialize_dof_vector()
    mg.vmult(z, _t(r))
    assert rel_err(z.cpu().numpy(), ref.vmult(r)) < 1e-9
    b = fine.assemble_rhs(1)
    bo = orc.LaplaceBK3(orc.mesh(8), 4, 5).rhs(1)
    assert rel_err(b.cpu().numpy(), bo) < 1e-12
    x = fine.initialize_dof_vector()
    its, r0, res = pmg.solve_cg(fine, b, x, pmg.PRECOND_MG, mg, 1e-10, 200)
    xr, rits, rr0, rres = ref.pcg(bo, precond=2, rel_tol=1e-10)
    assert abs(its - rits) <= 1, (its, rits)
    assert abs(r0 - rr0) <= 1e-12 * rr0 and res <= 1e-10 * r0
    assert rel_err(x.cpu().numpy(), xr) < 1e-9
    e_gpu = fine.l2_error(x, 1)
    e_ref = orc.LaplaceBK3(orc.mesh(8), 4, 5).l2_error(xr, 1)
    assert abs(e_gpu - e_ref) <= 1e-6 * e_ref  # both are discretisation errors of ~1e-7
    mg32, _, _, _, _ = pmg.geometric_hierarchy(pmg.box(8), 4, 3, 5, precision=pmg.FP32,
                                               outer_precision=pmg.FP64)
    x32 = fine.initialize_dof_vector()
    its32, _, res32 = pmg.solve_cg(fine, b, x32, pmg.PRECOND_MG, mg32, 1e-10, 200)
    assert res32 <= 1e-10 * r0 and abs(its32 - its) <= 1
    assert rel_err(x32.cpu().numpy(), xr) < 1e-8
    assert abs(fine.l2_error(x32, 1) - e_ref) <= 1e-6 * e_ref


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_config4_ph_hierarchy_6_3_1_matches_oracle(pmg, orc, prec):
    """ph-multigrid configured as the 48^3 Q6 workload, at 6^3 deformed
    cells: p 6 -> 3 -> 1 on the fine mesh, then Q1 h-levels 6^3 -> 3^3,
    nq = p + 1 per level; FP64 throughout, or FP32 V-cycle (smoother,
    transfers, levels) inside FP64 CG with the coarse solve in FP64.  The
    V-cycle (FP64), the iteration count (+-1) and the solution agree with the
    FP64 oracle; the mixed solve's distance to the oracle solution is the
    FP64-oracle error this configuration reports."""
    import torch
    print(f"[config4 6^3 Q6 6-3-1 {prec}]", flush=True)
    fine_box = pmg.box(6, deform=1, amplitude=0.05)
    levels = pmg.ph_levels(fine_box, 6, 2, p_sequence=[6, 3, 1])
    assert [(lv[0].nx, lv[1]) for lv in levels] == [(3, 1), (6, 1), (6, 3), (6, 6)]
    p = pmg.FP64 if prec == "fp64" else pmg.FP32
    mg, ops, _, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    m3, m6 = orc.mesh(3, deform=1, amplitude=0.05), orc.mesh(6, deform=1, amplitude=0.05)
    ref = orc.Multigrid([m3, m6, m6, m6], [1, 1, 3, 6], coarse_kind=1, coarse_tol=1e-12)
    if prec == "fp64":
        fine = ops[-1]
        r = seeded_uniform(fine.m(), 73)
        r[orc.boundary_dofs(m6, 6)] = 0
        z = fine.initialize_dof_vector()
        mg.vmult(z, _t(r))
        assert rel_err(z.cpu().numpy(), ref.vmult(r)) < 1e-9
    fine = pmg.LaplaceOperatorBK3(fine_box, 6)
    b = fine.assemble_rhs(0)
    bo = orc.LaplaceBK3(m6, 6).rhs(0)
    assert rel_err(b.cpu().numpy(), bo) < 1e-12
    x = fine.initialize_dof_vector()
    its, r0, res = pmg.solve_cg(fine, b, x, pmg.PRECOND_MG, mg, 1e-10, 200)
    xr, rits, rr0, _ = ref.pcg(bo, precond=2, rel_tol=1e-10)
    assert res <= 1e-10 * r0
    assert abs(its - rits) <= 1, (its, rits)
    err = rel_err(x.cpu().numpy(), xr)
    print(f"[config4 6^3 {prec}] its {its} (oracle {rits}), |u - u_oracle|/|u_oracle| = {err:.2e}",
          flush=True)
    assert err < (1e-9 if prec == "fp64" else 1e-8)
