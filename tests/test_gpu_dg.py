"""GPU parity of the SIP-DG operator, the DG<->CG transfer and DG multigrid
(libpmg.so) against the CPU oracle (pmo_dg.c).

Bars: FP64 vmult within 1e-12 relative (max-norm) of the oracle's pointwise
restatement of LaplaceOperatorDG::vmult; FP32 within 5e-6; penalties and the
cell diagonal within 1e-13 / 1e-12; transfers exact to 1e-15; MG-CG
iteration counts within +-1 of the oracle's DG hierarchy and solutions within
1e-9 relative.
"""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu

CASES = [
    # (cells (nx,ny,nz), degree, nq, deform, metric)
    ((3, 2, 2), 1, 2, 1, 0), ((2, 2, 3), 1, 3, 0, 0), ((3, 3, 2), 2, 3, 1, 0),
    ((2, 2, 2), 2, 4, 2, 0), ((4, 4, 4), 3, 4, 1, 0), ((3, 2, 3), 3, 5, 0, 0),
    ((2, 3, 2), 4, 5, 1, 0), ((2, 2, 2), 4, 6, 2, 1), ((2, 2, 2), 5, 6, 1, 0),
    ((2, 1, 2), 5, 7, 0, 0), ((2, 2, 1), 6, 7, 1, 0), ((1, 2, 2), 6, 8, 1, 1),
    ((3, 3, 3), 3, 4, 2, 1), ((3, 3, 3), 2, 3, 1, 1),
]
AMP = {0: 0.0, 1: 0.05, 2: 0.2}


def _t(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


def _pair(pmg, orc, cells, p, nq, deform, metric=0, precision=None, penalty_kind=0,
          neumann=0):
    nx, ny, nz = cells
    b = pmg.box(nx, ny, nz, deform=deform, amplitude=AMP[deform])
    m = orc.mesh(nx, ny, nz, deform=deform, amplitude=AMP[deform])
    op = pmg.LaplaceOperatorDG(b, p, nq, pmg.FP64 if precision is None else precision,
                               penalty_kind=penalty_kind, metric=metric, neumann_sides=neumann)
    ref = orc.LaplaceDG(m, p, nq, penalty_kind=penalty_kind, metric=metric,
                        neumann_mask=neumann)
    return op, ref


@pytest.mark.parametrize("cells,p,nq,deform,metric", CASES)
def test_dg_vmult_fp64_matches_oracle(pmg, orc, cells, p, nq, deform, metric):
    print(f"[dg fp64] cells={cells} p={p} nq={nq} deform={deform} metric={metric}", flush=True)
    op, ref = _pair(pmg, orc, cells, p, nq, deform, metric)
    x = seeded_uniform(op.m(), 3000 + 10 * p + nq)
    dst = op.initialize_dof_vector()
    op.vmult(dst, _t(x))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12


@pytest.mark.parametrize("cells,p,nq,deform", [((3, 2, 2), 1, 2, 1), ((3, 3, 3), 3, 4, 1),
                                               ((2, 2, 2), 6, 7, 0)])
def test_dg_vmult_fp32_matches_oracle(pmg, orc, cells, p, nq, deform):
    import torch
    print(f"[dg fp32] cells={cells} p={p}", flush=True)
    op, ref = _pair(pmg, orc, cells, p, nq, deform, precision=pmg.FP32)
    x = seeded_uniform(op.m(), 4000 + p)
    dst = op.initialize_dof_vector()
    op.vmult(dst, _t(x, torch.float32))
    assert rel_err(dst.cpu().numpy(), ref.vmult(x.astype(np.float32).astype(np.float64))) < 5e-6


def test_dg_penalty_kinds_and_neumann(pmg, orc):
    """Penalty kind 1 (2(p+1)^2/h on uniform cells) and Neumann box sides."""
    for kind, neumann in [(1, 0), (0, 0b010110), (1, 0b111111)]:
        print(f"[dg] penalty_kind={kind} neumann={neumann:06b}", flush=True)
        op, ref = _pair(pmg, orc, (3, 2, 2), 3, 4, 1, penalty_kind=kind, neumann=neumann)
        x = seeded_uniform(op.m(), 17 + kind)
        dst = op.initialize_dof_vector()
        op.vmult(dst, _t(x))
        assert rel_err(dst.cpu().numpy(), ref.vmult(x)) < 1e-12


def test_dg_penalty_values_match_oracle(pmg, orc):
    """The library's per-face penalty equals the oracle's per-cell-face one
    on both sides of every face."""
    op, ref = _pair(pmg, orc, (3, 2, 4), 3, 4, 1)
    sig = op.penalty()
    inner, bdry = pmg.box_dg_faces(op.box)
    inner, bdry = inner.reshape(-1, 4), bdry.reshape(-1, 3)
    s_ref = ref.sigma.reshape(-1, 6)
    F = len(inner)
    assert rel_err(sig[:F], s_ref[inner[:, 0], inner[:, 2]]) < 1e-13
    assert rel_err(sig[:F], s_ref[inner[:, 1], inner[:, 3]]) < 1e-13
    assert rel_err(sig[F:], s_ref[bdry[:, 0], bdry[:, 1]]) < 1e-13


def test_dg_apply_add_and_user_penalty(pmg, orc):
    op, ref = _pair(pmg, orc, (3, 3, 2), 2, 3, 1)
    x, base = seeded_uniform(op.m(), 5), seeded_uniform(op.m(), 6)
    dst = _t(base)
    op.apply_add(dst, _t(x))
    assert rel_err(dst.cpu().numpy(), base + ref.vmult(x)) < 1e-12
    # a caller-given penalty array replaces the computed one
    b = op.box
    pen = 1.7 * op.penalty()
    op2 = pmg.LaplaceOperatorDG(b, 2, 3, penalty=pen)
    inner, bdry = pmg.box_dg_faces(b)
    inner, bdry = inner.reshape(-1, 4), bdry.reshape(-1, 3)
    s = np.zeros((op.n_cells, 6))
    s[inner[:, 0], inner[:, 2]] = pen[:len(inner)]
    s[inner[:, 1], inner[:, 3]] = pen[:len(inner)]
    s[bdry[:, 0], bdry[:, 1]] = pen[len(inner):]
    ref2 = orc.LaplaceDG(ref.m, 2, 3, sigma=s.ravel())
    out = op2.initialize_dof_vector()
    op2.vmult(out, _t(x))
    assert rel_err(out.cpu().numpy(), ref2.vmult(x)) < 1e-12


def test_dg_cell_diagonal_matches_oracle(pmg, orc):
    op, ref = _pair(pmg, orc, (3, 2, 2), 3, 4, 1)
    op.compute_diagonal()
    inv = op.get_matrix_diagonal_inverse().cpu().numpy()
    assert rel_err(1.0 / inv, ref.cell_diagonal()) < 1e-12


def test_dg_symmetry_and_rhs_at_size(pmg, orc):
    """16^3 Q3 deformed: <Ax,y> = <x,Ay> to round-off; RHS against the oracle."""
    import torch
    b = pmg.box(16, deform=1, amplitude=0.05)
    op = pmg.LaplaceOperatorDG(b, 3)
    x = _t(seeded_uniform(op.m(), 21))
    y = _t(seeded_uniform(op.m(), 22))
    ax, ay = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(ax, x)
    op.vmult(ay, y)
    lhs, rhs = float(torch.dot(ax, y)), float(torch.dot(x, ay))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    m = orc.mesh(16, deform=1, amplitude=0.05)
    ref = orc.LaplaceDG(m, 3)
    assert rel_err(op.assemble_rhs(1).cpu().numpy(), ref.rhs(1)) < 1e-12


@pytest.mark.parametrize("p", [1, 3, 6])
def test_continuous_transfer_matches_oracle(pmg, orc, p):
    b = pmg.box(3, 2, 2)
    tr = pmg.ContinuousTransfer(b, p)
    cg_map = orc.dofmap(orc.mesh(3, 2, 2), p, True)
    ncg, ndg = pmg.box_n_dofs(b, p), pmg.box_n_cells(b) * (p + 1) ** 3
    xc, xf = seeded_uniform(ncg, 51), seeded_uniform(ndg, 52)
    yf0, yc0 = seeded_uniform(ndg, 53), seeded_uniform(ncg, 54)
    yf = _t(yf0)
    tr.prolongate_and_add(yf, _t(xc))
    assert rel_err(yf.cpu().numpy(), orc.dg_cg_prolongate(p, cg_map, xc, yf0.copy())) < 1e-15
    yc = _t(yc0)
    tr.restrict_and_add(yc, _t(xf))
    assert rel_err(yc.cpu().numpy(), orc.dg_cg_restrict(p, cg_map, xf, yc0.copy())) < 1e-14


@pytest.mark.parametrize("n,p,n_h", [(4, 3, 2), (4, 2, 3)])
def test_dg_mgcg_matches_oracle(pmg, orc, n, p, n_h):
    """The reference's DG hierarchy (Q1 h-levels, p-levels, DG(p) on top;
    tests/poisson_cube_bk3_fp64_dg/program.cc:294-460): V-cycle, MG-CG
    iterations and solution against the oracle; FP32 V-cycle in FP64 CG."""
    print(f"[dg mg] n={n} p={p} h-levels={n_h}", flush=True)
    fine = pmg.box(n, deform=1, amplitude=0.05)
    levels = pmg.ph_levels(fine, p, n_h, dg=True)
    mg, ops, trs, sms = pmg.build_hierarchy(levels)
    ms = [orc.mesh(lv[0].nx, deform=1, amplitude=0.05) for lv in levels]
    ref = orc.Multigrid(ms, [lv[1] for lv in levels], coarse_kind=1, coarse_tol=1e-12,
                        dg_top=True)
    top = ops[-1]
    r = seeded_uniform(top.m(), 91)
    z = top.initialize_dof_vector()
    mg.vmult(z, _t(r))
    assert rel_err(z.cpu().numpy(), ref.vmult(r)) < 1e-9
    b = top.assemble_rhs(0)
    x = top.initialize_dof_vector()
    its, r0, res = pmg.solve_cg(top, b, x, pmg.PRECOND_MG, mg, 1e-10, 200)
    xr, rits, _, _ = ref.pcg(b.cpu().numpy(), precond=2, rel_tol=1e-10)
    assert abs(its - rits) <= 1, (its, rits)
    assert res <= 1e-10 * r0
    assert rel_err(x.cpu().numpy(), xr) < 1e-9
    mg32, _, _, _ = pmg.build_hierarchy(levels, precision=pmg.FP32, outer_precision=pmg.FP64)
    x32 = top.initialize_dof_vector()
    its32, _, res32 = pmg.solve_cg(top, b, x32, pmg.PRECOND_MG, mg32, 1e-10, 200)
    assert res32 <= 1e-10 * r0 and its32 <= its + 2
    assert rel_err(x32.cpu().numpy(), x.cpu().numpy()) < 1e-8


def test_dg_discretisation_error_converges(pmg):
    """Manufactured u = sin sin sin on the unit cube (the deformations move
    boundary vertices, so u would not vanish on the boundary there): the DG
    Q2 L2 error falls by about 2^(p+1) per refinement."""
    errs = []
    for n in (4, 8):
        fine = pmg.box(n)
        mg, ops, _, _ = pmg.build_hierarchy(pmg.ph_levels(fine, 2, 2, dg=True))
        top = ops[-1]
        b = top.assemble_rhs(1)
        x = top.initialize_dof_vector()
        pmg.solve_cg(top, b, x, pmg.PRECOND_MG, mg, 1e-12, 200)
        errs.append(top.l2_error(x, 1))
    rate = np.log2(errs[0] / errs[1])
    assert rate > 2.6, errs


@pytest.mark.parametrize("cells,p,deform,neumann,kind,prec", [
    ((3, 3, 3), 2, 1, 0, 0, "fp64"), ((2, 3, 5), 3, 2, 0b101001, 1, "fp64"),
    ((1, 2, 3), 1, 0, 0b000110, 0, "fp64"), ((4, 1, 2), 2, 1, 0, 1, "fp32"),
    ((2, 2, 1), 4, 1, 0b111111, 0, "fp32")])
def test_dg_box_device_setup_matches_host_setup(pmg, cells, p, deform, neumann, kind, prec):
    """pmg_dg_operator_create_box (vertices, faces, penalty and geometry on
    the device) against the operator from host face lists: the same penalty
    per face (the same face numbering) and the same vmult and diagonal, to
    rounding."""
    import torch
    print(f"[dg box set-up] {cells} p={p} deform={deform} neumann={neumann} {prec}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    b = pmg.box(*cells, deform=deform, amplitude=AMP[deform])
    kw = dict(penalty_kind=kind, neumann_sides=neumann)
    d = pmg.LaplaceOperatorDG(b, p, None, precision, **kw)
    h = pmg.LaplaceOperatorDG(b, p, None, precision, device_setup=False, **kw)
    assert d.n_faces == h.n_faces and d.m() == h.m()
    assert rel_err(d.penalty(), h.penalty()) < 1e-14
    x = _t(seeded_uniform(d.m(), 231), dt)
    yd, yh = d.initialize_dof_vector(), h.initialize_dof_vector()
    d.vmult(yd, x)
    h.vmult(yh, x)
    assert rel_err(yd.double().cpu().numpy(), yh.double().cpu().numpy()) < (
        1e-13 if prec == "fp64" else 1e-6)
    d.compute_diagonal()
    h.compute_diagonal()
    # the device and host vertex generators round the deformation's sines
    # differently: the two operators agree to rounding, not bit for bit
    assert rel_err(d.get_matrix_diagonal_inverse().double().cpu().numpy(),
                   h.get_matrix_diagonal_inverse().double().cpu().numpy()) < 1e-13


def test_dg_rejects_bad_face_lists(pmg):
    b = pmg.box(2)
    cv = pmg.box_cell_vertices(b)
    inner, bdry = pmg.box_dg_faces(b)
    with pytest.raises(pmg.PMGError):  # a boundary face missing
        pmg.LaplaceOperatorDG.from_arrays(2, cv, inner, bdry[:-3])
    bad = inner.copy()
    bad[3] = bad[2]  # face+ must be face- ^ 1
    with pytest.raises(pmg.PMGError):
        pmg.LaplaceOperatorDG.from_arrays(2, cv, bad, bdry)
    with pytest.raises(pmg.PMGError):  # degree out of range
        pmg.LaplaceOperatorDG(b, 7)
