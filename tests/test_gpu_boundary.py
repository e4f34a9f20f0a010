"""The members of the reference's operator and transfer interfaces beyond
vmult / prolongate / restrict, through the C ABI:
  LaplaceOperatorBase::el and ::vmult_dummy
    (include/base/portable_laplace_operator_base.h:27-31,54; implementations
    portable_laplace_operator_bk3.h:233-357,586-595),
  MGTransferBase::reinit (include/base/portable_mg_transfer_base.h:32-35),
from Python and from the C++ mirror include/pmg.hpp (drivers/boundary_check)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR, rel_err, seeded_uniform

pytestmark = pytest.mark.gpu


def _t(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


@pytest.mark.parametrize("p,deform", [(2, 1), (4, 0)])
def test_el_is_the_assembled_diagonal(pmg, orc, p, deform):
    """el(i,i) = 1 / (inverse diagonal)_ii: the oracle's diagonal on free rows
    (pinned by dense element matrices), 1 on constrained rows; el(i,j != i)
    is refused as in the reference (Assert(row == col))."""
    print(f"[el] p={p} deform={deform}", flush=True)
    b = pmg.box(3, deform=deform, amplitude=0.05)
    m = orc.mesh(3, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p)
    op.compute_diagonal()
    ref = orc.LaplaceBK3(m, p).diagonal()
    cons = set(orc.boundary_dofs(m, p).tolist())
    rows = [0, op.m() // 2, op.m() - 1] + [i for i in range(op.m()) if i not in cons][:40:7]
    for i in rows:
        v = op.el(i, i)
        if i in cons:
            assert v == 1.0
        else:
            assert abs(v - ref[i]) <= 1e-12 * abs(ref[i]), i
    with pytest.raises(pmg.PMGError):
        op.el(0, 1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vmult_dummy_combinations(pmg, orc, prec):
    """One GPU, no ghosts: (ghost, compute) = (1,1) vmult; (0,1) the cell
    loop alone (constrained rows 0); (1,0) only dst_c = src_c; (0,0) zero --
    each into a dst holding stale values."""
    import torch
    print(f"[vmult_dummy] {prec}", flush=True)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    dt = torch.float64 if prec == "fp64" else torch.float32
    tol = 1e-12 if prec == "fp64" else 2e-6
    b, m, p = pmg.box(3, deform=1, amplitude=0.05), orc.mesh(3, deform=1, amplitude=0.05), 3
    op = pmg.LaplaceOperatorBK3(b, p, precision=precision)
    x = seeded_uniform(op.m(), 131)
    ax = orc.LaplaceBK3(m, p).vmult(x)
    cons = orc.boundary_dofs(m, p)
    cell, copy = ax.copy(), np.zeros(op.m())
    cell[cons] = 0.0
    copy[cons] = x[cons]
    expect = {(1, 1): ax, (0, 1): cell, (1, 0): copy, (0, 0): np.zeros(op.m())}
    for (g, c), e in expect.items():
        y = _t(seeded_uniform(op.m(), 132), dt)
        op.vmult_dummy(y, _t(x, dt), g, c)
        got = y.double().cpu().numpy()
        if not e.any():
            assert not got.any()
        else:
            assert np.abs(got - e).max() <= tol * np.abs(ax).max(), (g, c)
    # DG: no constrained rows, so (1,0) and (0,0) are zero and (0,1) is vmult
    dg = pmg.LaplaceOperatorDG(pmg.box(2, deform=1, amplitude=0.05), 2, precision=precision)
    xd = _t(seeded_uniform(dg.m(), 133), dt)
    yv = dg.initialize_dof_vector()
    dg.vmult(yv, xd)
    for g, c in [(0, 1), (1, 1), (1, 0)]:
        y = _t(seeded_uniform(dg.m(), 134), dt)
        dg.vmult_dummy(y, xd, g, c)
        if c:
            assert rel_err(y.double().cpu().numpy(), yv.double().cpu().numpy()) < tol, (g, c)
        else:
            assert not y.any()


@pytest.mark.parametrize("kind", ["h", "p", "dg"])
def test_transfer_reinit_equals_a_fresh_transfer(pmg, kind):
    """A transfer reinit'ed to another box (other sizes, maps, constraints)
    is the transfer built for that box."""
    import torch
    print(f"[transfer reinit] {kind}", flush=True)
    a, b = pmg.box(2, deform=1, amplitude=0.05), pmg.box(3, deform=1, amplitude=0.05)
    if kind == "h":
        tr, fresh = pmg.GeometricTransfer(a, 2), pmg.GeometricTransfer(b, 2)
        mc, mf, _ = pmg.box_h_transfer_maps(b, 2, weights=False)
        cons = pmg.box_boundary_dofs(pmg.refine(b), 2)
        nc, nf = pmg.box_n_dofs(b, 2), pmg.box_n_dofs(pmg.refine(b), 2)
    elif kind == "p":
        tr, fresh = pmg.PolynomialTransfer(a, 1, 3), pmg.PolynomialTransfer(b, 1, 3)
        mc, mf, _ = pmg.box_p_transfer_maps(b, 1, 3, weights=False)
        cons = pmg.box_boundary_dofs(b, 3)
        nc, nf = pmg.box_n_dofs(b, 1), pmg.box_n_dofs(b, 3)
    else:
        tr, fresh = pmg.ContinuousTransfer(a, 2), pmg.ContinuousTransfer(b, 2)
        mc, mf, cons = pmg.box_dofmap(b, 2, True), None, None
        nc, nf = pmg.box_n_dofs(b, 2), pmg.box_n_cells(b) * 27
    tr.reinit(mc, mf, cons, nc, nf)
    assert tr.info()[:4] == fresh.info()[:4]
    xc, xf = _t(seeded_uniform(nc, 141)), _t(seeded_uniform(nf, 142))
    y1, y2 = _t(np.zeros(nf)), _t(np.zeros(nf))
    tr.prolongate_and_add(y1, xc)
    fresh.prolongate_and_add(y2, xc)
    assert rel_err(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14
    r1, r2 = _t(np.zeros(nc)), _t(np.zeros(nc))
    tr.restrict_and_add(r1, xf)
    fresh.restrict_and_add(r2, xf)
    assert rel_err(r1.cpu().numpy(), r2.cpu().numpy()) < 1e-14


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vcycle_after_transfer_reinit(pmg, monkeypatch, prec):
    """A V-cycle whose transfer is rebuilt after its graph was captured gives
    the eager V-cycle's result over the same rebuild: first the fine
    h-transfer with every coarse entry constrained (no coarse correction
    reaches the fine level), then rebuilt back to its own maps (the first
    result again)."""
    import torch
    print(f"[vcycle after reinit] {prec}", flush=True)
    p = pmg.FP64 if prec == "fp64" else pmg.FP32
    levels = [(pmg.box(2), 2), (pmg.box(4), 2), (pmg.box(8), 2)]
    mg_g, ops_g, trs_g, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    monkeypatch.setenv("PMG_VCYCLE_GRAPH", "0")
    mg_e, ops_e, trs_e, _ = pmg.build_hierarchy(levels, precision=p, outer_precision=pmg.FP64)
    assert mg_g.uses_graph() and not mg_e.uses_graph()
    n = ops_g[-1].m()
    r = _t(seeded_uniform(n, 151))
    z0 = torch.zeros_like(r)
    mg_g.vmult(z0, r)
    torch.cuda.synchronize()
    mc, mf, _ = pmg.box_h_transfer_maps(pmg.box(4), 2, weights=False)
    cons = pmg.box_boundary_dofs(pmg.box(8), 2)
    nc = pmg.box_n_dofs(pmg.box(4), 2)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for coarse_map, changed in ((np.full_like(mc, 0xFFFFFFFF), True), (mc, False)):
        for trs in (trs_g, trs_e):
            trs[-1].reinit(coarse_map, mf, cons, nc, n)
        zg, ze = torch.zeros_like(r), torch.zeros_like(r)
        mg_g.vmult(zg, r)
        mg_e.vmult(ze, r)
        torch.cuda.synchronize()
        assert mg_g.uses_graph()
        assert rel_err(zg.cpu().numpy(), ze.cpu().numpy()) < tol, changed
        d0 = rel_err(zg.cpu().numpy(), z0.cpu().numpy())
        assert (d0 > 1e-3) if changed else (d0 < tol), (changed, d0)


def test_cpp_mirror_boundary_members():
    """drivers/boundary_check (C++ on include/pmg.hpp) calls el, vmult_dummy,
    MGTransferBase::reinit and the affine-geometry option and reports their
    deviations."""
    exe = os.path.join(PKG_DIR, "drivers", "boundary_check")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    print("[cpp boundary_check]", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    r = {k: float(v) for k, v in (it.split("=") for it in out.stdout.split())}
    assert r["el"] < 1e-14
    for k in ("dummy_11", "dummy_01", "dummy_10", "dummy_00"):
        assert r[k] < 1e-14, k
    assert r["reinit_p"] < 1e-14 and r["reinit_r"] < 1e-14
    assert r["affine"] < 1e-13 and r["affine_refused"] == 1
