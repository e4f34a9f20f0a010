"""Degenerate and extreme inputs: no cells, every DoF constrained, a single
Neumann cell, and a mesh whose geometric factors exceed 2^31 entries."""
import numpy as np
import pytest

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu


def test_operator_without_cells(pmg):
    """C = 0: vmult returns 0 on free DoFs and copies constrained ones;
    apply_add adds nothing; the diagonal has no cell to come from (error)."""
    import torch
    op = pmg.LaplaceOperatorBK3.from_arrays(2, np.zeros(0, dtype=np.uint32), 5,
                                            np.array([3], dtype=np.uint32), G=np.zeros(0),
                                            n_q_points_1d=3)
    assert op.m() == 5
    x = torch.tensor(seeded_uniform(5, 601), device="cuda")
    y = op.initialize_dof_vector()
    op.vmult(y, x)
    expect = np.zeros(5)
    expect[3] = x[3].item()
    assert np.array_equal(y.cpu().numpy(), expect)
    base = torch.tensor(seeded_uniform(5, 602), device="cuda")
    z = base.clone()
    op.apply_add(z, x)
    assert torch.equal(z, base)
    with pytest.raises(pmg.PMGError):
        op.compute_diagonal()


def test_every_dof_constrained(pmg):
    """One Q1 cell with homogeneous Dirichlet everywhere: the operator is the
    identity, its inverse diagonal 1."""
    import torch
    b = pmg.box(1)
    op = pmg.LaplaceOperatorBK3(b, 1)
    assert len(pmg.box_boundary_dofs(b, 1)) == op.m() == 8
    x = torch.tensor(seeded_uniform(8, 603), device="cuda")
    y = op.initialize_dof_vector()
    op.vmult(y, x)
    assert torch.equal(y, x)
    op.compute_diagonal()
    assert torch.equal(op.get_matrix_diagonal_inverse(), torch.ones(8, dtype=torch.float64,
                                                                    device="cuda"))


@pytest.mark.parametrize("p", [1, 3, 6])
def test_single_neumann_cell(pmg, orc, p):
    """One deformed cell without constraints: symmetric, A 1 = 0, and equal to
    the oracle's cell matrix applied to x."""
    import torch
    b = pmg.box(1, deform=2, amplitude=0.1)
    n = pmg.box_n_dofs(b, p)
    op = pmg.LaplaceOperatorBK3.from_arrays(p, pmg.box_dofmap(b, p, False), n,
                                            np.zeros(0, dtype=np.uint32),
                                            cell_vertices=pmg.box_cell_vertices(b))
    x, w = seeded_uniform(n, 604), seeded_uniform(n, 605)
    ax, aw, a1 = (op.initialize_dof_vector() for _ in range(3))
    op.vmult(ax, torch.tensor(x, device="cuda"))
    op.vmult(aw, torch.tensor(w, device="cuda"))
    op.vmult(a1, torch.ones(n, dtype=torch.float64, device="cuda"))
    lhs, rhs = float(np.dot(ax.cpu().numpy(), w)), float(np.dot(x, aw.cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert a1.abs().max().item() <= 1e-12 * ax.abs().max().item()
    ref = orc.LaplaceBK3(orc.mesh(1, deform=2, amplitude=0.1), p, p + 1)
    ref.map = ref.plain_map  # no Dirichlet rows: the cell operator alone
    assert rel_err(ax.cpu().numpy(), ref.cells_add(x, np.zeros(n))) < 1e-12


def test_geometry_beyond_2e31_entries(pmg):
    """160^3 cells at Q4 (nq = 5): G holds 3.07e9 values (> 2^31, 24.6 GB) and
    N = 641^3 = 263 374 721 DoFs.  Size-independent checks at full size:
    symmetry <Ax,y> = <x,Ay> and A 1 = 0 on rows away from the boundary."""
    import torch
    n, p = 160, 4
    b = pmg.box(n)
    op = pmg.LaplaceOperatorBK3(b, p, 5, pmg.FP64)
    N = op.m()
    assert N == (4 * n + 1) ** 3 and op.n_cells * 6 * 125 > 2 ** 31
    g = torch.Generator(device="cuda").manual_seed(607)
    x = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) - 0.5
    cons = torch.tensor(pmg.box_boundary_dofs(b, p).astype(np.int64), device="cuda")
    x[cons] = 0
    y = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) - 0.5
    y[cons] = 0
    ax = op.initialize_dof_vector()
    op.vmult(ax, x)
    lhs = torch.dot(ax, y).item()
    del ax
    ay = op.initialize_dof_vector()
    op.vmult(ay, y)
    rhs = torch.dot(x, ay).item()
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    del ay, x, y
    a1 = op.initialize_dof_vector()
    op.vmult(a1, torch.ones(N, dtype=torch.float64, device="cuda"))
    M = 4 * n + 1
    inner = a1.reshape(M, M, M)[5:-5, 5:-5, 5:-5]
    assert inner.abs().max().item() < 1e-11 * a1.abs().max().item()


def test_constrained_dofs_are_eliminated_from_the_map(pmg):
    """A map that still names the constrained DoFs (no 0xFFFFFFFF entries)
    gives the operator of the map that marks them invalid: constrained DoFs
    are eliminated from rows and columns, and their rows are the identity."""
    import torch
    b, p = pmg.box(3, deform=1, amplitude=0.05), 2
    n = pmg.box_n_dofs(b, p)
    cons = pmg.box_boundary_dofs(b, p)
    op_plain = pmg.LaplaceOperatorBK3.from_arrays(p, pmg.box_dofmap(b, p, False), n, cons,
                                                  cell_vertices=pmg.box_cell_vertices(b))
    op = pmg.LaplaceOperatorBK3(b, p)
    x = torch.tensor(seeded_uniform(n, 608), device="cuda")
    y1, y2 = op.initialize_dof_vector(), op_plain.initialize_dof_vector()
    op.vmult(y1, x)
    op_plain.vmult(y2, x)
    assert rel_err(y2.cpu().numpy(), y1.cpu().numpy()) < 1e-14
    assert torch.equal(y2[torch.tensor(cons.astype(np.int64), device="cuda")],
                       x[torch.tensor(cons.astype(np.int64), device="cuda")])
