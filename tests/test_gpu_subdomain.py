"""GPU parity of the domain-decomposition subdomain operator against the oracle.

SubdomainLaplaceOperator (include/operators/portable_subdomain_laplace_operator.h):
vmult_bk3 is the operator's vmult with every subdomain constraint eliminated;
vmult_neumann and vmult_interface_cell_range use the plain index set, in which
only the physical boundary DoFs are eliminated, over every cell or over the
interface cells.  The subdomain is a box whose x = hi face is the interface.
Bar: FP64 within 1e-12 relative (max-norm) of the oracle; FP32 within 2e-6.
"""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform, subdomain_sets

pytestmark = pytest.mark.gpu

CASES = [
    # (nx, ny, nz, degree, nq, deform)
    (3, 3, 3, 1, 2, 0), (3, 3, 3, 2, 3, 1), (4, 3, 2, 3, 4, 1), (3, 3, 3, 3, 5, 0),
    (5, 4, 4, 4, 5, 1), (3, 3, 3, 4, 6, 0), (2, 2, 2, 5, 6, 1), (2, 3, 2, 6, 7, 1),
]


def _setup(pmg, orc, nx, ny, nz, p, nq, deform, precision):
    b = pmg.box(nx, ny, nz, deform=deform, amplitude=0.05)
    m = orc.mesh(nx, ny, nz, deform=deform, amplitude=0.05)
    op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
    ref = orc.LaplaceBK3(m, p, nq)
    phys, iface, icells = subdomain_sets(nx, ny, nz, p)
    op.set_subdomain(pmg.box_dofmap(b, p, False), phys, icells)
    return op, ref, phys, icells


def _apply(op, fn, x, dtype):
    import torch
    dst = op.initialize_dof_vector()
    fn(dst, torch.tensor(x, dtype=dtype, device="cuda"))
    return dst.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("nx,ny,nz,p,nq,deform", CASES)
def test_subdomain_vmults_fp64_match_oracle(pmg, orc, nx, ny, nz, p, nq, deform):
    import torch
    print(f"[subdomain fp64] {nx}x{ny}x{nz} p={p} nq={nq} deform={deform}", flush=True)
    op, ref, phys, icells = _setup(pmg, orc, nx, ny, nz, p, nq, deform, pmg.FP64)
    x = seeded_uniform(op.m(), 300 + 10 * p + nq)
    assert rel_err(_apply(op, op.vmult, x, torch.float64), ref.vmult(x)) < 1e-12
    y_neu = ref.vmult_subdomain(x, phys)
    assert rel_err(_apply(op, op.vmult_neumann, x, torch.float64), y_neu) < 1e-12
    y_int = ref.vmult_subdomain(x, phys, icells)
    assert rel_err(_apply(op, op.vmult_interface_cell_range, x, torch.float64), y_int) < 1e-12


@pytest.mark.parametrize("nx,ny,nz,p,nq,deform", [(3, 3, 3, 2, 3, 1), (5, 4, 4, 4, 5, 1),
                                                  (2, 3, 2, 6, 7, 1)])
def test_subdomain_vmults_fp32_match_oracle(pmg, orc, nx, ny, nz, p, nq, deform):
    import torch
    print(f"[subdomain fp32] {nx}x{ny}x{nz} p={p} nq={nq}", flush=True)
    op, ref, phys, icells = _setup(pmg, orc, nx, ny, nz, p, nq, deform, pmg.FP32)
    x = seeded_uniform(op.m(), 400 + p).astype(np.float32).astype(np.float64)
    assert rel_err(_apply(op, op.vmult_neumann, x, torch.float32),
                   ref.vmult_subdomain(x, phys)) < 2e-6
    assert rel_err(_apply(op, op.vmult_interface_cell_range, x, torch.float32),
                   ref.vmult_subdomain(x, phys, icells)) < 2e-6


def test_subdomain_with_permuted_cells_and_edge_cases(pmg, orc):
    """Caller cell order shuffled (interface cells named in that order); an
    empty interface; no physical DoFs (pure Neumann: constants in the kernel)."""
    import torch
    nx, p, nq = 4, 3, 4
    print(f"[subdomain permuted] {nx}^3 p={p} nq={nq}", flush=True)
    b = pmg.box(nx, deform=1, amplitude=0.05)
    ref = orc.LaplaceBK3(orc.mesh(nx, deform=1, amplitude=0.05), p, nq)
    nm3, N = (p + 1) ** 3, pmg.box_n_dofs(b, p)
    C = pmg.box_n_cells(b)
    perm = np.random.default_rng(5).permutation(C)
    inv = np.empty(C, dtype=np.int64)
    inv[perm] = np.arange(C)
    cmap = pmg.box_dofmap(b, p, True).reshape(C, nm3)[perm].ravel()
    pmap = pmg.box_dofmap(b, p, False).reshape(C, nm3)[perm].ravel()
    cv = pmg.box_cell_vertices(b).reshape(C, -1)[perm].ravel()
    op = pmg.LaplaceOperatorBK3.from_arrays(p, cmap, N, pmg.box_boundary_dofs(b, p),
                                            cell_vertices=cv, n_q_points_1d=nq)
    phys, iface, icells = subdomain_sets(nx, nx, nx, p)
    op.set_subdomain(pmap, phys, inv[icells].astype(np.uint32))
    x = seeded_uniform(N, 77)
    assert rel_err(_apply(op, op.vmult_neumann, x, torch.float64),
                   ref.vmult_subdomain(x, phys)) < 1e-12
    assert rel_err(_apply(op, op.vmult_interface_cell_range, x, torch.float64),
                   ref.vmult_subdomain(x, phys, icells)) < 1e-12
    # no interface cells: zero on the free rows, src on the physical ones
    op.set_subdomain(pmap, phys, np.zeros(0, np.uint32))
    y = _apply(op, op.vmult_interface_cell_range, x, torch.float64)
    free = np.setdiff1d(np.arange(N), phys)
    assert np.all(y[free] == 0.0) and np.array_equal(y[phys], x[phys])
    # no physical DoFs: the constant vector is in the kernel
    op.set_subdomain(pmap, np.zeros(0, np.uint32), inv[icells].astype(np.uint32))
    y = _apply(op, op.vmult_neumann, np.ones(N), torch.float64)
    assert np.abs(y).max() < 1e-12 * np.abs(ref.vmult_subdomain(x, phys)).max()


def test_subdomain_errors(pmg):
    import torch
    b = pmg.box(2)
    op = pmg.LaplaceOperatorBK3(b, 2)
    x = op.initialize_dof_vector()
    y = op.initialize_dof_vector()
    with pytest.raises(pmg.PMGError, match="set_subdomain"):
        op.vmult_neumann(y, x)
    pmap = pmg.box_dofmap(b, 2, False)
    phys, _, icells = subdomain_sets(2, 2, 2, 2)
    with pytest.raises(pmg.PMGError, match="out of range"):
        op.set_subdomain(pmap, phys, np.array([0, 8], np.uint32))
    with pytest.raises(pmg.PMGError, match="twice"):
        op.set_subdomain(pmap, phys, np.array([1, 1], np.uint32))
    with pytest.raises(pmg.PMGError, match="out of range"):
        op.set_subdomain(pmap, np.array([op.m()], np.uint32), icells)
    op.set_subdomain(pmap, phys, icells)
    with pytest.raises(pmg.PMGError, match="alias"):
        op.vmult_interface_cell_range(x, x)
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_subdomain_neumann_on_patch_schedule(pmg, orc, prec):
    """The plain index set gets the main map's scatter schedule: the patch
    schedule when the operator uses it (forced here, and the FP32 default on
    more than 4096 cells)."""
    import torch
    precision = pmg.FP64 if prec == "fp64" else pmg.FP32
    dt, tol = (torch.float64, 1e-12) if prec == "fp64" else (torch.float32, 2e-6)
    for (n, p, nq, force) in [(4, 3, 4, True), (17, 2, 3, prec == "fp64")]:
        print(f"[subdomain patch {prec}] {n}^3 p={p} nq={nq} forced={force}", flush=True)
        b = pmg.box(n, deform=1, amplitude=0.05)
        ref = orc.LaplaceBK3(orc.mesh(n, deform=1, amplitude=0.05), p, nq)
        op = pmg.LaplaceOperatorBK3(b, p, nq, precision)
        if force:
            op.set_schedule(pmg.SCHEDULE_PATCH, -1)
        phys, _, icells = subdomain_sets(n, n, n, p)
        op.set_subdomain(pmg.box_dofmap(b, p, False), phys, icells)
        x = seeded_uniform(op.m(), 500 + n)
        if prec == "fp32":
            x = x.astype(np.float32).astype(np.float64)
        assert rel_err(_apply(op, op.vmult_neumann, x, dt), ref.vmult_subdomain(x, phys)) < tol
        assert rel_err(_apply(op, op.vmult_interface_cell_range, x, dt),
                       ref.vmult_subdomain(x, phys, icells)) < tol
