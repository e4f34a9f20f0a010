"""Index data built on the device (pmg_operator_create_box,
pmg_transfer_create_box, and the device elimination / compression / patch
build behind pmg_operator_create) against host builds of the same arrays:

  * the box maps, boundary lists and vertices against the host mesh
    generator (pmg_box_*), which the oracle tests pin to the reference's
    numbering (tests/poisson_cube_bk3/program.cc:281-566);
  * the constrained-DoF elimination (setup_dof_indices_per_color,
    include/operators/portable_laplace_operator_bk3.h:360-437), the x-line
    compression and the patch schedule against numpy restatements, on box maps
    and on shuffled numberings with extra constraints and ragged last batches;
  * box transfers against transfers from host maps and constraint lists
    (GeometricTransfer::setup_dof_indices / setup_weights,
    include/multigrid/portable_geometric_transfer.h:1344-1486).

Index data is compared bit for bit; vertices from the device sine within
2 ulp-scale (1e-15 absolute on the unit box)."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu

INV = 0xFFFFFFFF


def _t(a, dtype=None):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda")


# ---------------------------------------------------------------- host restatements
def eliminate_np(dofmap, cons):
    m = dofmap.copy()
    mask = np.zeros(int(max(dofmap[dofmap != INV].max(initial=0), cons.max(initial=0))) + 1, bool)
    mask[cons] = True
    valid = m != INV
    hit = np.zeros_like(valid)
    hit[valid] = mask[m[valid]]
    m[hit] = INV
    return m


def lines_np(imap, n_cells, nm, bound, invalid):
    """x-line starts of the cells whose nm^2 x-lines are runs of nm consecutive
    valid indices below bound (zeros elsewhere), and those cells' flags; None
    when fewer than half of the cells qualify (the arrays are not kept)."""
    m = imap.reshape(n_cells, nm * nm, nm).astype(np.int64)
    l0 = m[:, :, 0]
    run = np.all(m == l0[:, :, None] + np.arange(nm), axis=2)
    ok = ((l0 != invalid) & (l0 + nm <= bound) & run).all(axis=1)
    if 2 * int(ok.sum()) < n_cells:
        return None, None
    return np.where(ok[:, None], l0, 0).reshape(-1), ok.astype(np.uint8)


def patches_np(imap, n_cells, nm3, E, n_dofs):
    """Per batch of E cells: the ascending DoFs of its valid entries (bit 31:
    the DoF has entries outside the batch), their count, and every entry's slot."""
    valid_all = imap != INV
    mult = np.bincount(imap[valid_all].astype(np.int64), minlength=n_dofs)
    nb, pmax = -(-n_cells // E), E * nm3
    pdof = np.zeros(nb * pmax, np.uint32)
    pcount = np.zeros(nb, np.uint32)
    pidx = np.full(n_cells * nm3, 0xFFFF, np.uint16)
    for b in range(nb):
        e0, e1 = b * E * nm3, min(n_cells, (b + 1) * E) * nm3
        ent = imap[e0:e1]
        valid = ent != INV
        dofs, inv, cnt = np.unique(ent[valid], return_inverse=True, return_counts=True)
        pidx[e0 + np.nonzero(valid)[0]] = inv
        shared = (cnt < mult[dofs]).astype(np.uint32) << 31
        pdof[b * pmax:b * pmax + len(dofs)] = dofs.astype(np.uint32) | shared
        pcount[b] = len(dofs)
    return pdof, pcount, pidx


def batch_cells(nq):
    return max(256 // (nq * nq), 1)


def colors_np(imap, n_cells, nm3, E):
    """Greedy colouring of each batch's cells in cell order (the smallest
    colour no earlier cell of the batch sharing a DoF has) and the colour
    count per batch, 0 where a cell names one DoF twice."""
    nb = -(-n_cells // E)
    pcolor = np.zeros(n_cells, np.uint8)
    pncol = np.zeros(nb, np.uint8)
    for b in range(nb):
        masks, dup, ncol = {}, False, 0
        for c in range(b * E, min(n_cells, (b + 1) * E)):
            ent = imap[c * nm3:(c + 1) * nm3]
            ent = [int(d) for d in ent if d != INV]
            dup = dup or len(set(ent)) != len(ent)
            forbidden = 0
            for d in set(ent):
                forbidden |= masks.get(d, 0)
            col = 0
            while (forbidden >> col) & 1:
                col += 1
            for d in set(ent):
                masks[d] = masks.get(d, 0) | (1 << col)
            pcolor[c] = col
            ncol = max(ncol, col + 1)
        pncol[b] = 0 if dup else ncol
    return pcolor, pncol


def chunks_np(pdof, pcount, pmax):
    """Run records of every 64 patch slots of every batch: first DoFs of runs
    0..3 of consecutive DoFs among the valid slots (bit 31 cleared), their
    first slots s1..s3 (64 if absent) and the number of runs (0 if more than
    4) packed in word 4."""
    nb, qc = len(pcount), (pmax + 63) // 64
    pd = (pdof.reshape(nb, pmax) & 0x7FFFFFFF).astype(np.int64)
    rec = np.zeros((nb * qc, 8), np.uint32)
    for b in range(nb):
        for q in range(qc):
            d = pd[b, q * 64:min(q * 64 + 64, int(pcount[b]))]
            r = rec[b * qc + q]
            starts = [0] + [u for u in range(1, len(d)) if d[u] != d[u - 1] + 1] if len(d) else []
            use = starts[:4]  # a chunk of more runs keeps its first 4 and says 0 runs
            s = use + [64] * (4 - len(use))
            if not use:
                s[0] = 0
            for k, u in enumerate(use):
                r[k] = d[u]
            nr = 1 if not use else (len(starts) if len(starts) <= 4 else 0)
            r[4] = s[1] | (s[2] << 8) | (s[3] << 16) | (nr << 24)
    return rec


def check_operator_index_data(pmg, op, dofmap, cons, p, nq, n_dofs):
    """op's device arrays against the restatements over the caller's map."""
    n_cells = len(dofmap) // (p + 1) ** 3
    nm, nm3 = p + 1, (p + 1) ** 3
    imap = eliminate_np(dofmap, cons)
    assert np.array_equal(op.index_data("map"), imap)
    assert np.array_equal(op.index_data("cell_order"), np.arange(n_cells, dtype=np.uint32))
    lines, flags = lines_np(imap, n_cells, nm, n_dofs, INV)
    if lines is None:
        assert op.index_data("lines").size == 0 and op.index_data("line_flags").size == 0
    else:
        assert np.array_equal(op.index_data("lines"), lines)
        assert np.array_equal(op.index_data("line_flags"), flags)
    op.set_schedule(pmg.SCHEDULE_PATCH)
    pdof, pcount, pidx = patches_np(imap, n_cells, nm3, batch_cells(nq), n_dofs)
    assert np.array_equal(op.index_data("patch_counts"), pcount)
    assert np.array_equal(op.index_data("patch_dofs"), pdof)
    assert np.array_equal(op.index_data("patch_slots"), pidx)
    pc, pn = colors_np(imap, n_cells, nm3, batch_cells(nq))
    assert np.array_equal(op.index_data("patch_colors"), pc)
    assert np.array_equal(op.index_data("patch_ncolors"), pn)
    ch = chunks_np(pdof, pcount, batch_cells(nq) * nm3)
    if 2 * int((ch[:, 4] >> 24 != 0).sum()) >= len(ch):
        assert np.array_equal(op.index_data("patch_chunks").reshape(-1, 8), ch)
    else:
        assert op.index_data("patch_chunks").size == 0
    pl, pf = lines_np(pidx.astype(np.uint32), n_cells, nm, 0xFFFF, 0xFFFF)
    if pl is None:
        assert op.index_data("patch_lines").size == 0
    else:
        assert np.array_equal(op.index_data("patch_lines"), pl.astype(np.uint16))
        assert np.array_equal(op.index_data("patch_line_flags"), pf)


# ---------------------------------------------------------------- operator
@pytest.mark.parametrize("n,p,nq,deform", [(3, 1, 2, 1), (3, 4, 5, 1), (4, 2, 4, 2),
                                           ((5, 3, 2), 3, 4, 0), ((2, 3, 7), 6, 7, 1),
                                           (6, 5, 6, 2)])
def test_box_operator_index_data_matches_host_build(pmg, n, p, nq, deform):
    """pmg_operator_create_box: map, constraints, lines, patches and vertices
    equal the host mesh generator's arrays and the restatements over them
    (ragged last batches: 27 cells in batches of 10, 30 in 7, ...)."""
    dims = n if isinstance(n, tuple) else (n, n, n)
    print(f"[device set-up] box {dims} p={p} nq={nq} deform={deform}", flush=True)
    b = pmg.box(*dims, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq)
    n_dofs = pmg.box_n_dofs(b, p)
    dofmap = pmg.box_dofmap(b, p, True)
    cons = pmg.box_boundary_dofs(b, p)
    assert op.m() == n_dofs
    assert np.array_equal(op.index_data("constrained"), cons)
    assert np.array_equal(op.index_data("map"), dofmap)
    check_operator_index_data(pmg, op, dofmap, cons, p, nq, n_dofs)
    cv = pmg.box_cell_vertices(b)
    dv = op.index_data("vertices")
    if deform == 1:
        assert np.abs(dv - cv).max() <= 1e-15
    else:
        assert np.array_equal(dv, cv)


@pytest.mark.parametrize("p,nq,n_extra", [(2, 3, 0), (3, 5, 40), (1, 3, 7)])
def test_array_operator_device_elimination_on_shuffled_numbering(pmg, p, nq, n_extra):
    """pmg_operator_create on a shuffled DoF numbering with extra constrained
    DoFs: the device elimination, compression and patches equal the
    restatements (no x-line runs: the lines are dropped)."""
    print(f"[device set-up] shuffled p={p} nq={nq} extra={n_extra}", flush=True)
    b = pmg.box(3, 3, 4, deform=1, amplitude=0.05)
    n_dofs = pmg.box_n_dofs(b, p)
    rng = np.random.Generator(np.random.PCG64(97))
    relabel = rng.permutation(n_dofs).astype(np.uint32)
    plain = pmg.box_dofmap(b, p, False)
    dofmap = relabel[plain]
    cons = np.union1d(relabel[pmg.box_boundary_dofs(b, p)],
                      rng.choice(n_dofs, n_extra, replace=False)).astype(np.uint32)
    op = pmg.LaplaceOperatorBK3.from_arrays(p, dofmap, n_dofs, cons,
                                            cell_vertices=pmg.box_cell_vertices(b),
                                            n_q_points_1d=nq)
    assert np.array_equal(op.index_data("constrained"), cons)
    check_operator_index_data(pmg, op, dofmap, cons, p, nq, n_dofs)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("accum", ["atomic", "colored"])
@pytest.mark.parametrize("p,nq,dims", [(4, 5, (12, 3, 3)), (6, 7, (5, 3, 2)), (3, 4, (20, 2, 3)),
                                       (1, 2, (4, 4, 5))])
def test_patch_accumulation_modes_match_oracle(pmg, orc, monkeypatch, prec, accum, p, nq, dims):
    """The patch schedule with its cells summed into the LDS patch by LDS
    atomics or colour by colour gives the oracle's vmult."""
    import torch
    print(f"[patch accumulation] {accum} {prec} p={p} nq={nq} {dims}", flush=True)
    monkeypatch.setenv("PMG_PATCH_ACCUM", accum)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    tol = 1e-12 if prec == "fp64" else 2e-6
    b = pmg.box(*dims, deform=1, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
    op.set_schedule(pmg.SCHEDULE_PATCH)
    assert op.index_data("patch_ncolors").min() >= 1
    ref = orc.LaplaceBK3(orc.mesh(*dims, deform=1, amplitude=0.05), p, nq)
    x = seeded_uniform(op.m(), 231)
    y = op.initialize_dof_vector()
    op.vmult(y, _t(x, op.dtype))
    assert rel_err(y.double().cpu().numpy(), ref.vmult(x)) < tol


def test_patch_colouring_falls_back_on_repeated_dofs(pmg, monkeypatch):
    """A cell that names one DoF twice keeps LDS atomics for its batch
    (pncol 0); the vmult still equals the atomic schedule's."""
    import torch
    monkeypatch.setenv("PMG_PATCH_ACCUM", "colored")
    b, p, nq = pmg.box(12, 3, 3, deform=1, amplitude=0.05), 2, 3
    n_dofs = pmg.box_n_dofs(b, p)
    dofmap = pmg.box_dofmap(b, p, True)
    nm3 = (p + 1) ** 3
    c = 37  # an interior cell: its entries 13 and 14 are both valid
    assert dofmap[c * nm3 + 13] != INV and dofmap[c * nm3 + 14] != INV
    dofmap[c * nm3 + 14] = dofmap[c * nm3 + 13]
    op = pmg.LaplaceOperatorBK3.from_arrays(p, dofmap, n_dofs, pmg.box_boundary_dofs(b, p),
                                            cell_vertices=pmg.box_cell_vertices(b),
                                            n_q_points_1d=nq)
    op.set_schedule(pmg.SCHEDULE_PATCH)
    pn = op.index_data("patch_ncolors")
    E = batch_cells(nq)
    assert pn[c // E] == 0 and (np.delete(pn, c // E) > 0).all()
    x = _t(seeded_uniform(n_dofs, 241))
    y1, y2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.vmult(y1, x)
    op.set_schedule(pmg.SCHEDULE_ATOMIC)
    op.vmult(y2, x)
    assert rel_err(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14


def test_device_range_checks_name_the_entry(pmg):
    """Out-of-range map entries and constrained DoFs are refused by the
    device checks, naming the first offending map position."""
    b = pmg.box(2)
    p = 2
    n_dofs = pmg.box_n_dofs(b, p)
    dofmap = pmg.box_dofmap(b, p, True)
    cons = pmg.box_boundary_dofs(b, p)
    cv = pmg.box_cell_vertices(b)
    bad = dofmap.copy()
    bad[[41, 77]] = n_dofs
    with pytest.raises(pmg.PMGError, match="out of range at 41"):
        pmg.LaplaceOperatorBK3.from_arrays(p, bad, n_dofs, cons, cell_vertices=cv)
    with pytest.raises(pmg.PMGError, match="constrained DoF out of range"):
        pmg.LaplaceOperatorBK3.from_arrays(p, dofmap, n_dofs, np.append(cons, n_dofs + 3),
                                           cell_vertices=cv)
    op = pmg.LaplaceOperatorBK3(b, p)
    with pytest.raises(pmg.PMGError, match="unknown index array"):
        pmg._index_data(pmg.lib().pmg_operator_index_data, op.h, (99, np.uint32))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("deform", [1, 2])
def test_box_operator_vmult_equals_array_operator(pmg, prec, deform):
    """The box-built and the array-built operator apply the same matrix: bit
    for bit with the deterministic gather schedule where the vertices agree
    bit for bit (shear), to round-off with the device sine."""
    import torch
    print(f"[device set-up] vmult {prec} deform={deform}", flush=True)
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    b = pmg.box(5, 4, 6, deform=deform, amplitude=0.05)
    p = 3
    a = pmg.LaplaceOperatorBK3.from_arrays(p, pmg.box_dofmap(b, p, True), pmg.box_n_dofs(b, p),
                                           pmg.box_boundary_dofs(b, p),
                                           cell_vertices=pmg.box_cell_vertices(b),
                                           precision=precision)
    d = pmg.LaplaceOperatorBK3(b, p, precision=precision)
    for op in (a, d):
        op.set_schedule(pmg.SCHEDULE_GATHER)
    x = _t(seeded_uniform(d.m(), 201), dt)
    ya, yd = a.initialize_dof_vector(), d.initialize_dof_vector()
    a.vmult(ya, x)
    d.vmult(yd, x)
    if deform == 2:
        assert torch.equal(ya, yd)
    else:
        assert rel_err(yd.double().cpu().numpy(), ya.double().cpu().numpy()) < (
            1e-14 if prec == "fp64" else 1e-6)
    # the box operator declares its structure: the grid schedule is available
    d.set_schedule(pmg.SCHEDULE_GRID)
    yg = d.initialize_dof_vector()
    d.vmult(yg, x)
    assert rel_err(yg.double().cpu().numpy(), yd.double().cpu().numpy()) < (
        1e-14 if prec == "fp64" else 1e-6)


def test_subdomain_device_set_up_matches_restatement(pmg):
    """set_subdomain builds the plain map on the device: the physical DoFs
    eliminated from the caller's map (checked through the Neumann vmult
    against the operator of the same plain map)."""
    import torch
    from conftest import subdomain_sets
    print("[device set-up] subdomain", flush=True)
    b, p = pmg.box(3), 2
    phys, _, icells = subdomain_sets(3, 3, 3, p)
    pmap = pmg.box_dofmap(b, p, False)
    op = pmg.LaplaceOperatorBK3(b, p)
    op.set_subdomain(pmap, phys, icells)
    ref = pmg.LaplaceOperatorBK3.from_arrays(p, pmap, pmg.box_n_dofs(b, p), phys,
                                             cell_vertices=pmg.box_cell_vertices(b))
    x = _t(seeded_uniform(op.m(), 211))
    y1, y2 = op.initialize_dof_vector(), ref.initialize_dof_vector()
    op.vmult_neumann(y1, x)
    ref.vmult(y2, x)
    assert rel_err(y1.cpu().numpy(), y2.cpu().numpy()) < 1e-14


# ---------------------------------------------------------------- transfers
@pytest.mark.parametrize("kind,p,pf,n", [("h", 1, 1, 3), ("h", 4, 4, 3), ("h", 2, 2, (2, 3, 5)),
                                         ("p", 1, 4, 4), ("p", 3, 6, 3), ("dg", 2, 2, 3)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_box_transfer_index_data_matches_host_build(pmg, monkeypatch, kind, p, pf, n, prec):
    """pmg_transfer_create_box against the transfer from host maps and the
    fine constraint list: the same coarse map, owner-marked fine map, fine
    x-line words and restriction inverse maps, bit for bit; the same
    prolongation and restriction."""
    import torch
    dims = n if isinstance(n, tuple) else (n, n, n)
    print(f"[device set-up] transfer {kind} p={p} pf={pf} box={dims} {prec}", flush=True)
    monkeypatch.setenv("PMG_RESTRICT_GATHER", "1")  # build the restriction inverse map too
    precision, dt = (pmg.FP64, torch.float64) if prec == "fp64" else (pmg.FP32, torch.float32)
    b = pmg.box(*dims, deform=1, amplitude=0.05)
    if kind == "h":
        d = pmg.GeometricTransfer(b, p, precision, device_weights=True)
        mc, mf, _ = pmg.box_h_transfer_maps(b, p, weights=False)
        cons = pmg.box_boundary_dofs(pmg.refine(b), p)
        a = pmg._Transfer(pmg.TRANSFER_H, precision, p, p, mc, mf, None, d.n_c, d.n_f, 1.0,
                          cons=cons)
        names = ["coarse_map", "fine_map", "fine_lines", "restrict_offsets", "restrict_positions"]
    elif kind == "p":
        d = pmg.PolynomialTransfer(b, p, pf, precision, device_weights=True)
        mc, mf, _ = pmg.box_p_transfer_maps(b, p, pf, weights=False)
        cons = pmg.box_boundary_dofs(b, pf)
        a = pmg._Transfer(pmg.TRANSFER_P, precision, p, pf, mc, mf, None, d.n_c, d.n_f, 1.0,
                          cons=cons)
        names = ["coarse_map", "fine_map", "fine_lines", "restrict_offsets", "restrict_positions"]
    else:
        d = pmg.ContinuousTransfer(b, p, precision)
        a = pmg._Transfer(pmg.TRANSFER_DG_CG, precision, p, p, pmg.box_dofmap(b, p, True), None,
                          None, d.n_c, d.n_f, 1.0)
        names = ["coarse_map", "dg_offsets", "dg_list"]
    assert d.info() == a.info()
    for name in names:
        ad, dd = a.index_data(name), d.index_data(name)
        assert ad.size > 0, name
        assert np.array_equal(dd, ad), name
    if kind != "dg":
        off, pos = restrict_buffer_layout(d.index_data("coarse_map"), d.n_c)
        assert np.array_equal(d.index_data("restrict_offsets"), off)
        assert np.array_equal(d.index_data("restrict_positions"), pos)
    xc, xf = _t(seeded_uniform(d.n_c, 221), dt), _t(seeded_uniform(d.n_f, 222), dt)
    ya, yd = (torch.zeros(d.n_f, dtype=dt, device="cuda") for _ in range(2))
    a.prolongate_and_add(ya, xc)
    d.prolongate_and_add(yd, xc)
    assert torch.equal(ya, yd)
    ra, rd = (torch.zeros(d.n_c, dtype=dt, device="cuda") for _ in range(2))
    a.restrict_and_add(ra, xf)
    d.restrict_and_add(rd, xf)
    assert torch.equal(ra, rd)  # the gather restriction sums in a fixed order


def restrict_buffer_layout(mc, n_c):
    """The restriction buffer's layout restated: coarse DoF i owns the range
    [off[i], off[i+1]), its entries ordered by their index in the coarse map;
    pos[e] is entry e's place (invalid for constrained entries)."""
    mc = np.asarray(mc, dtype=np.uint32)
    valid = mc != np.uint32(0xFFFFFFFF)
    off = np.zeros(n_c + 1, dtype=np.uint32)
    off[1:] = np.cumsum(np.bincount(mc[valid].astype(np.int64), minlength=n_c))
    ent = np.flatnonzero(valid)
    order = ent[np.argsort(mc[ent], kind="stable")]
    pos = np.full(mc.size, 0xFFFFFFFF, dtype=np.uint32)
    pos[order] = np.arange(order.size, dtype=np.uint32)
    return off, pos


def test_box_transfer_rejects_bad_arguments(pmg):
    import ctypes
    h = ctypes.c_void_p()
    lib = pmg.lib()
    b = pmg.box(2)
    assert lib.pmg_transfer_create_box(pmg.TRANSFER_H, pmg.FP64, ctypes.byref(b), 2, 3, 1.0,
                                       ctypes.byref(h)) != pmg.OK  # h needs pc == pf
    assert lib.pmg_transfer_create_box(pmg.TRANSFER_P, pmg.FP64, ctypes.byref(b), 3, 2, 1.0,
                                       ctypes.byref(h)) != pmg.OK  # p needs pc < pf
    assert lib.pmg_transfer_create_box(pmg.TRANSFER_P, pmg.FP64, None, 1, 2, 1.0,
                                       ctypes.byref(h)) != pmg.OK
    bad = pmg.box(2)
    bad.nx = 0
    assert lib.pmg_operator_create_box(ctypes.byref(bad), 2, 3, pmg.FP64, ctypes.byref(h)) != pmg.OK
    assert lib.pmg_operator_create_box(ctypes.byref(b), 7, 8, pmg.FP64, ctypes.byref(h)) != pmg.OK


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_patch_chunk_runs_with_fallback_chunks(pmg, prec):
    """A box numbering with a few DoFs relabelled: most 64-slot chunks of the
    patch lists are <= 4 runs, some are not and load their entries; the patch
    vmult equals the atomic schedule's on the same operator."""
    import torch
    print(f"[patch chunks] partial relabel {prec}", flush=True)
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    p, nq, dims = 4, 5, (12, 3, 4)
    b = pmg.box(*dims, deform=1, amplitude=0.05)
    n_dofs = pmg.box_n_dofs(b, p)
    rng = np.random.Generator(np.random.PCG64(1234))
    relabel = np.arange(n_dofs, dtype=np.uint32)
    pick = rng.choice(n_dofs, n_dofs // 100, replace=False)
    relabel[pick] = relabel[rng.permutation(pick)]
    dofmap = relabel[pmg.box_dofmap(b, p, False)]
    cons = np.sort(relabel[pmg.box_boundary_dofs(b, p)]).astype(np.uint32)
    op = pmg.LaplaceOperatorBK3.from_arrays(p, dofmap, n_dofs, cons,
                                            cell_vertices=pmg.box_cell_vertices(b),
                                            n_q_points_1d=nq, precision=precision)
    op.set_schedule(pmg.SCHEDULE_PATCH)
    ch = op.index_data("patch_chunks").reshape(-1, 8)
    nr = ch[:, 4] >> 24
    assert ch.size > 0 and (nr == 0).any() and (nr > 0).sum() * 2 >= len(nr)
    x = _t(seeded_uniform(n_dofs, 77), op.dtype)
    y_patch = op.initialize_dof_vector()
    op.vmult(y_patch, x)
    op.set_schedule(pmg.SCHEDULE_ATOMIC)
    y_atomic = op.initialize_dof_vector()
    op.vmult(y_atomic, x)
    tol = 1e-12 if prec == "fp64" else 2e-6
    assert rel_err(y_patch.double().cpu().numpy(), y_atomic.double().cpu().numpy()) < tol
