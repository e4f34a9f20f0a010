"""Arbitrary numberings.  deal.II's DoFHandler numbers DoFs cell by cell and
hands over cells in tree order, not lexicographically as the box helpers do.
The operators, transfers and the MG-CG solve must be equivariant under any
permutation pi of the global DoF indices and sigma of the cell order: on
permuted inputs the GPU returns the permuted result of the lexicographic box
inputs.  This exercises the map-driven parts that a lexicographic numbering
leaves easy: the patch schedule's batch patches and ownership, the transfers'
owner entries, the inverse DG->CG map and the DG face lists.
"""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF


def _perm(n, seed):
    return np.random.Generator(np.random.PCG64(seed)).permutation(n).astype(np.int64)


def _permute_map(m, nm3, pi, sigma):
    """Rows (cells) in the order sigma, entries renumbered by pi, invalid kept."""
    m = m.reshape(-1, nm3)[sigma]
    out = m.copy()
    ok = m != INVALID
    out[ok] = pi[m[ok]]
    return np.ascontiguousarray(out.ravel(), dtype=np.uint32)


def _bk3_pair(pmg, b, p, nq, prec, seed):
    N, C, nm3 = pmg.box_n_dofs(b, p), pmg.box_n_cells(b), (p + 1) ** 3
    pi, sigma = _perm(N, seed), _perm(C, seed + 1)
    m = _permute_map(pmg.box_dofmap(b, p), nm3, pi, sigma)
    cv = pmg.box_cell_vertices(b).reshape(C, 24)[sigma].ravel()
    cons = pi[pmg.box_boundary_dofs(b, p).astype(np.int64)].astype(np.uint32)
    op2 = pmg.LaplaceOperatorBK3.from_arrays(p, m, N, cons, cell_vertices=cv, n_q_points_1d=nq,
                                             precision=prec)
    return pmg.LaplaceOperatorBK3(b, p, nq, prec), op2, pi


def _to(pi, v):
    """v in the box numbering -> the permuted numbering (out[pi[i]] = v[i])."""
    out = np.empty_like(v)
    out[pi] = v
    return out


@pytest.mark.parametrize("p,nq,prec", [(1, 2, 0), (2, 3, 0), (3, 4, 0), (4, 5, 0), (4, 5, 1),
                                       (3, 4, 1), (6, 7, 0)])
def test_bk3_equivariant_under_renumbering(pmg, p, nq, prec):
    import torch
    print(f"[numbering bk3] p={p} nq={nq} prec={prec}", flush=True)
    b = pmg.box(3, deform=1, amplitude=0.05)
    op1, op2, pi = _bk3_pair(pmg, b, p, nq, prec, 500 + p)
    dt = torch.float64 if prec == pmg.FP64 else torch.float32
    tol = 1e-12 if prec == pmg.FP64 else 2e-6
    x = seeded_uniform(op1.m(), 510 + p)
    y1 = op1.initialize_dof_vector()
    op1.vmult(y1, torch.tensor(x, dtype=dt, device="cuda"))
    y1 = y1.cpu().numpy().astype(np.float64)
    x2 = torch.tensor(_to(pi, x), dtype=dt, device="cuda")
    for sched in (None, pmg.SCHEDULE_ATOMIC, pmg.SCHEDULE_COLORED, pmg.SCHEDULE_PATCH,
                  pmg.SCHEDULE_GATHER):
        if sched is not None:
            op2.set_schedule(sched)
        y2 = op2.initialize_dof_vector()
        op2.vmult(y2, x2)
        assert rel_err(y2.cpu().numpy().astype(np.float64)[pi], y1) < tol, sched
    op1.compute_diagonal()
    op2.compute_diagonal()
    d1 = op1.get_matrix_diagonal_inverse().cpu().numpy().astype(np.float64)
    d2 = op2.get_matrix_diagonal_inverse().cpu().numpy().astype(np.float64)
    assert rel_err(d2[pi], d1) < tol
    if prec == pmg.FP64:
        r1, r2 = op1.assemble_rhs(1).cpu().numpy(), op2.assemble_rhs(1).cpu().numpy()
        assert rel_err(r2[pi], r1) < 1e-13
        e1 = op1.l2_error(torch.tensor(x, device="cuda"), 1)
        e2 = op2.l2_error(x2, 1)
        assert abs(e1 - e2) <= 1e-13 * abs(e1)


@pytest.mark.parametrize("p,nq,prec", [(3, 4, 0), (4, 5, 0), (4, 5, 1), (2, 3, 1)])
def test_patch_schedule_with_partly_scattered_numbering(pmg, p, nq, prec):
    """DoFs of the bottom quarter of the box shuffled, the rest lexicographic:
    batches there have hundreds of runs of consecutive DoFs, the others a few
    dozen, so one patch launch mixes batches read through run tables with
    batches read through their plain DoF lists."""
    import torch
    print(f"[numbering patch mixed] p={p} nq={nq} prec={prec}", flush=True)
    b = pmg.box(8, deform=1, amplitude=0.05)
    N, nm3 = pmg.box_n_dofs(b, p), (p + 1) ** 3
    pi = np.arange(N, dtype=np.int64)
    q = N // 4
    pi[:q] = _perm(q, 520 + p)
    m = _permute_map(pmg.box_dofmap(b, p), nm3, pi, np.arange(pmg.box_n_cells(b)))
    cons = pi[pmg.box_boundary_dofs(b, p).astype(np.int64)].astype(np.uint32)
    op1 = pmg.LaplaceOperatorBK3(b, p, nq, prec)
    op2 = pmg.LaplaceOperatorBK3.from_arrays(p, m, N, cons, cell_vertices=pmg.box_cell_vertices(b),
                                             n_q_points_1d=nq, precision=prec)
    op1.set_schedule(pmg.SCHEDULE_ATOMIC)
    op2.set_schedule(pmg.SCHEDULE_PATCH)
    dt = torch.float64 if prec == pmg.FP64 else torch.float32
    tol = 1e-12 if prec == pmg.FP64 else 2e-6
    x = seeded_uniform(N, 530 + p)
    y1, y2 = op1.initialize_dof_vector(), op2.initialize_dof_vector()
    op1.vmult(y1, torch.tensor(x, dtype=dt, device="cuda"))
    op2.vmult(y2, torch.tensor(_to(pi, x), dtype=dt, device="cuda"))
    y1 = y1.cpu().numpy().astype(np.float64)
    assert rel_err(y2.cpu().numpy().astype(np.float64)[pi], y1) < tol
    base = seeded_uniform(N, 540 + p)
    z2 = torch.tensor(_to(pi, base), dtype=dt, device="cuda")
    op2.apply_add(z2, torch.tensor(_to(pi, x), dtype=dt, device="cuda"))
    free = np.ones(N, dtype=bool)
    free[pmg.box_boundary_dofs(b, p).astype(np.int64)] = False
    assert rel_err(z2.cpu().numpy().astype(np.float64)[pi][free], (base + y1)[free]) < tol


@pytest.mark.parametrize("kind", ["h", "p"])
def test_transfers_equivariant_under_renumbering(pmg, kind):
    import torch
    if kind == "h":
        bc, pc, pf = pmg.box(2, deform=1, amplitude=0.05), 2, 2
        mc, mf, w = pmg.box_h_transfer_maps(bc, pc)
        nc, nf = pmg.box_n_dofs(bc, pc), pmg.box_n_dofs(pmg.refine(bc), pf)
        nmf3 = (2 * pc + 1) ** 3
        tr1 = pmg.GeometricTransfer(bc, pc)
        k = pmg.TRANSFER_H
    else:
        bc, pc, pf = pmg.box(3, deform=1, amplitude=0.05), 1, 3
        mc, mf, w = pmg.box_p_transfer_maps(bc, pc, pf)
        nc, nf = pmg.box_n_dofs(bc, pc), pmg.box_n_dofs(bc, pf)
        nmf3 = (pf + 1) ** 3
        tr1 = pmg.PolynomialTransfer(bc, pc, pf)
        k = pmg.TRANSFER_P
    C = pmg.box_n_cells(bc)
    pic, pif, rho = _perm(nc, 520), _perm(nf, 521), _perm(C, 522)
    mc2 = _permute_map(mc, (pc + 1) ** 3, pic, rho)
    mf2 = _permute_map(mf, nmf3, pif, rho)
    w2 = w.reshape(C, nmf3)[rho].ravel()
    tr2 = pmg._Transfer(k, pmg.FP64, pc, pf, mc2, mf2, w2, nc, nf, 1.0)
    # the same transfer with its weights built on the device from the fine
    # constraints: on a shuffled fine numbering the x-line words do not apply
    # and the owner marks stay a full map
    fine_box = pmg.refine(bc) if kind == "h" else bc
    cons2 = pif[pmg.box_boundary_dofs(fine_box, pf).astype(np.int64)].astype(np.uint32)
    tr3 = pmg._Transfer(k, pmg.FP64, pc, pf, mc2, mf2, None, nc, nf, 1.0, cons=cons2)
    assert tr1.layout() == pmg.TRANSFER_LAYOUT_OWNER_LINES
    assert tr2.layout() == pmg.TRANSFER_LAYOUT_OWNER_MAP
    assert tr3.layout() == pmg.TRANSFER_LAYOUT_OWNER_MAP
    xc, yf0 = seeded_uniform(nc, 523), seeded_uniform(nf, 524)
    yf1 = torch.tensor(yf0, device="cuda")
    tr1.prolongate_and_add(yf1, torch.tensor(xc, device="cuda"))
    for tr in (tr2, tr3):
        yf2 = torch.tensor(_to(pif, yf0), device="cuda")
        tr.prolongate_and_add(yf2, torch.tensor(_to(pic, xc), device="cuda"))
        assert rel_err(yf2.cpu().numpy()[pif], yf1.cpu().numpy()) < 1e-13
    xf, yc0 = seeded_uniform(nf, 525), seeded_uniform(nc, 526)
    yc1 = torch.tensor(yc0, device="cuda")
    tr1.restrict_and_add(yc1, torch.tensor(xf, device="cuda"))
    for tr in (tr2, tr3):
        yc2 = torch.tensor(_to(pic, yc0), device="cuda")
        tr.restrict_and_add(yc2, torch.tensor(_to(pif, xf), device="cuda"))
        assert rel_err(yc2.cpu().numpy()[pic], yc1.cpu().numpy()) < 1e-13


def _dg_pair(pmg, b, p, nq, seed):
    """Box DG operator and the same operator on permuted cells and shuffled face lists."""
    C, nm3 = pmg.box_n_cells(b), (p + 1) ** 3
    sigma = _perm(C, seed)
    inv = np.empty(C, dtype=np.int64)
    inv[sigma] = np.arange(C)
    inner, bdry = pmg.box_dg_faces(b)
    inner = inner.reshape(-1, 4).astype(np.int64)
    bdry = bdry.reshape(-1, 3).astype(np.int64)
    inner2 = inner.copy()
    inner2[:, 0], inner2[:, 1] = inv[inner[:, 0]], inv[inner[:, 1]]
    inner2 = inner2[_perm(len(inner2), seed + 1)]
    bdry2 = bdry.copy()
    bdry2[:, 0] = inv[bdry[:, 0]]
    bdry2 = bdry2[_perm(len(bdry2), seed + 2)]
    cv = pmg.box_cell_vertices(b).reshape(C, 24)[sigma].ravel()
    op2 = pmg.LaplaceOperatorDG.from_arrays(p, cv, inner2.astype(np.uint32).ravel(),
                                            bdry2.astype(np.uint32).ravel(), n_q_points_1d=nq,
                                            penalty_kind=1)
    op1 = pmg.LaplaceOperatorDG(b, p, nq, pmg.FP64, penalty_kind=1)
    # DG DoFs are cell-major: DoF (r, l) of the permuted operator is (sigma[r], l)
    pi = (inv[:, None] * nm3 + np.arange(nm3)[None, :]).ravel()  # box DoF (c, l) -> (inv[c], l)
    return op1, op2, pi, sigma


def test_dg_equivariant_under_cell_and_face_order(pmg):
    import torch
    b = pmg.box(3, deform=1, amplitude=0.05)
    op1, op2, pi, _ = _dg_pair(pmg, b, 2, 3, 530)
    x = seeded_uniform(op1.m(), 531)
    y1 = op1.initialize_dof_vector()
    op1.vmult(y1, torch.tensor(x, device="cuda"))
    y2 = op2.initialize_dof_vector()
    op2.vmult(y2, torch.tensor(_to(pi, x), device="cuda"))
    assert rel_err(y2.cpu().numpy()[pi], y1.cpu().numpy()) < 1e-12


def test_dg_cg_transfer_equivariant(pmg):
    import torch
    b, p = pmg.box(3, deform=1, amplitude=0.05), 2
    N, C, nm3 = pmg.box_n_dofs(b, p), pmg.box_n_cells(b), (p + 1) ** 3
    pic, sigma = _perm(N, 540), _perm(C, 541)
    inv = np.empty(C, dtype=np.int64)
    inv[sigma] = np.arange(C)
    mc2 = _permute_map(pmg.box_dofmap(b, p, True), nm3, pic, sigma)
    tr1 = pmg.ContinuousTransfer(b, p)
    tr2 = pmg._Transfer(pmg.TRANSFER_DG_CG, pmg.FP64, p, p, mc2, None, None, N, C * nm3, 1.0)
    pid = (inv[:, None] * nm3 + np.arange(nm3)[None, :]).ravel()
    xc, yd0 = seeded_uniform(N, 542), seeded_uniform(C * nm3, 543)
    yd1 = torch.tensor(yd0, device="cuda")
    tr1.prolongate_and_add(yd1, torch.tensor(xc, device="cuda"))
    yd2 = torch.tensor(_to(pid, yd0), device="cuda")
    tr2.prolongate_and_add(yd2, torch.tensor(_to(pic, xc), device="cuda"))
    assert rel_err(yd2.cpu().numpy()[pid], yd1.cpu().numpy()) < 1e-14
    xd, yc0 = seeded_uniform(C * nm3, 544), seeded_uniform(N, 545)
    yc1 = torch.tensor(yc0, device="cuda")
    tr1.restrict_and_add(yc1, torch.tensor(xd, device="cuda"))
    yc2 = torch.tensor(_to(pic, yc0), device="cuda")
    tr2.restrict_and_add(yc2, torch.tensor(_to(pid, xd), device="cuda"))
    assert rel_err(yc2.cpu().numpy()[pic], yc1.cpu().numpy()) < 1e-14


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mg_cg_on_renumbered_hierarchy(pmg, prec):
    """h-MG Q2 over 2^3 -> 4^3 -> 8^3 cells with every level, and each transfer's
    cell order, renumbered independently: same iteration count (+-1: the
    eigenvalue estimate starts from (i mod 11) in DoF order, as deal.II's) and
    the permuted solution of the box hierarchy."""
    import torch
    P = pmg.FP64 if prec == "fp64" else pmg.FP32
    p, boxes = 2, pmg.coarsen_boxes(pmg.box(8, deform=1, amplitude=0.05), 3)
    mg1, ops1, _, _ = pmg.build_hierarchy([(bb, p) for bb in boxes], precision=P,
                                          outer_precision=pmg.FP64)
    ops2, trs2, sms2, pis = [], [None], [None], []
    for l, bb in enumerate(boxes):
        _, op, pi = _bk3_pair(pmg, bb, p, p + 1, P, 550 + 10 * l)
        op.compute_diagonal()
        ops2.append(op)
        pis.append(pi)
        if l == 0:
            continue
        bc = boxes[l - 1]
        mc, mf, w = pmg.box_h_transfer_maps(bc, p)
        C = pmg.box_n_cells(bc)
        rho = _perm(C, 551 + 10 * l)
        trs2.append(pmg._Transfer(pmg.TRANSFER_H, P, p, p,
                                  _permute_map(mc, (p + 1) ** 3, pis[l - 1], rho),
                                  _permute_map(mf, (2 * p + 1) ** 3, pi, rho),
                                  w.reshape(C, -1)[rho].ravel(), ops2[l - 1].m(), op.m(), 1.0))
        sms2.append(pmg.PreconditionChebyshev(op, 3, 15.0, 10))
    mg2 = pmg.VCycleMultigrid(ops2, trs2, sms2, pmg.COARSE_PCG, 1e-12, pmg.FP64)
    fine1 = pmg.LaplaceOperatorBK3(boxes[-1], p, p + 1, pmg.FP64)
    fine2 = _bk3_pair(pmg, boxes[-1], p, p + 1, pmg.FP64, 550 + 10 * (len(boxes) - 1))[1]
    pi = pis[-1]
    b1, b2 = fine1.assemble_rhs(1), fine2.assemble_rhs(1)
    x1, x2 = fine1.initialize_dof_vector(), fine2.initialize_dof_vector()
    it1, _, _ = pmg.solve_cg(fine1, b1, x1, pmg.PRECOND_MG, mg1, 1e-10, 100)
    it2, _, _ = pmg.solve_cg(fine2, b2, x2, pmg.PRECOND_MG, mg2, 1e-10, 100)
    assert abs(it1 - it2) <= 1, (it1, it2)
    assert rel_err(x2.cpu().numpy()[pi], x1.cpu().numpy()) < 1e-7
    assert abs(fine1.l2_error(x1, 1) - fine2.l2_error(x2, 1)) <= 1e-6 * fine1.l2_error(x1, 1)
    torch.cuda.synchronize()
