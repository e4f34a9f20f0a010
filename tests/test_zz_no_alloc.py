"""pmg_cg_solve_noalloc (CG on the library's fixed device work array) against
pmg_cg_solve on the same operator, right-hand side and preconditioner: two
calls in a row, the second on a smaller problem, so the second call runs on
what the first left in the shared array.  Both go through the C ABI."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import rel_err, seeded_uniform

pytestmark = pytest.mark.gpu

_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
FP64_MAX_DOFS, FP32_MAX_DOFS = 2096096, 4192192


def _noalloc(pmg):
    L = pmg.lib()
    f = L.pmg_cg_solve_noalloc
    P = ctypes.POINTER
    f.restype = _i
    f.argtypes = [_vp, _i, _vp, _vp, _vp, _d, _i, P(_i), P(_d), P(_d), _vp]
    return f


def _solve(pmg, fn, op, b, x, precond, mg, rel_tol, max_it):
    its, r0, r = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    st = fn(op.h, precond, mg.h if mg is not None else None, b.data_ptr(), x.data_ptr(), rel_tol,
            max_it, ctypes.byref(its), ctypes.byref(r0), ctypes.byref(r),
            torch.cuda.current_stream().cuda_stream)
    return st, its.value, r0.value, r.value


def _case(pmg, fn_new, fine, degree, n_levels, prec, precond, seed):
    print(f"[noalloc] box {fine.nx} Q{degree} prec {prec} precond {precond}", flush=True)
    dt = torch.float64 if prec == pmg.FP64 else torch.float32
    if precond == pmg.PRECOND_MG:
        mg, ops, _, _, _ = pmg.geometric_hierarchy(fine, degree, n_levels, precision=prec)
        op = ops[-1]
    else:
        mg, op = None, pmg.LaplaceOperatorBK3(fine, degree, None, prec)
        op.compute_diagonal()
    b = torch.tensor(seeded_uniform(op.m(), seed), dtype=dt, device="cuda")
    tol = 1e-10 if prec == pmg.FP64 else 1e-5
    x_ref = torch.zeros_like(b)
    st, its_ref, r0_ref, r_ref = _solve(pmg, pmg.lib().pmg_cg_solve, op, b, x_ref, precond, mg,
                                        tol, 2000)
    assert st == pmg.OK, pmg.lib().pmg_last_error()
    x = torch.zeros_like(b)
    st, its, r0, r = _solve(pmg, fn_new, op, b, x, precond, mg, tol, 2000)
    assert st == pmg.OK, pmg.lib().pmg_last_error()
    torch.cuda.synchronize()
    # same iteration (device FP atomics may round the last bits differently)
    assert abs(its - its_ref) <= 1, (its, its_ref)
    assert abs(r0 - r0_ref) <= 1e-12 * r0_ref
    assert r <= tol * r0
    bar = 1e-9 if prec == pmg.FP64 else 1e-4
    assert rel_err(x.cpu().numpy(), x_ref.cpu().numpy()) < bar
    return op.m()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_noalloc_cg_matches_cg_on_two_calls(pmg, prec, precond):
    p = pmg.FP64 if prec == "fp64" else pmg.FP32
    pc = pmg.PRECOND_JACOBI if precond == "jacobi" else pmg.PRECOND_MG
    fn = _noalloc(pmg)
    n1 = _case(pmg, fn, pmg.box(8, deform=1, amplitude=0.05), 3, 3, p, pc, 901)
    n2 = _case(pmg, fn, pmg.box(4), 2, 2, p, pc, 902)  # smaller: stale contents above it
    assert n2 < n1


def test_noalloc_cg_rejects_sizes_beyond_the_array(pmg):
    """A problem above the fixed array's capacity is refused with a clear
    error, before any device work (32^3 Q4 FP64: 2 146 689 DoFs)."""
    print("[noalloc] size limit", flush=True)
    fn = _noalloc(pmg)
    op = pmg.LaplaceOperatorBK3(pmg.box(32), 4, None, pmg.FP64)
    assert op.m() > FP64_MAX_DOFS
    op.compute_diagonal()
    b = op.initialize_dof_vector()
    x = op.initialize_dof_vector()
    st, _, _, _ = _solve(pmg, fn, op, b, x, pmg.PRECOND_JACOBI, None, 1e-8, 10)
    assert st == pmg.ERR_SIZE
    msg = pmg.lib().pmg_last_error().decode()
    assert "exceed the fixed work array" in msg and str(FP64_MAX_DOFS) in msg
    # the same operator in FP32 fits (twice the DoFs per byte)
    op32 = pmg.LaplaceOperatorBK3(pmg.box(32), 4, None, pmg.FP32)
    op32.compute_diagonal()
    assert op32.m() <= FP32_MAX_DOFS
    b32 = torch.tensor(seeded_uniform(op32.m(), 903), dtype=torch.float32, device="cuda")
    x32 = op32.initialize_dof_vector()
    st, its, r0, r = _solve(pmg, fn, op32, b32, x32, pmg.PRECOND_JACOBI, None, 1e-3, 5000)
    assert st == pmg.OK, pmg.lib().pmg_last_error()
    assert r <= 1e-3 * r0 and its > 0
