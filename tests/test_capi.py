"""CPU-side checks of the C ABI: libpmg.so loads, exports every entry point that
include/pmg.h declares, and its host-only mesh helpers reproduce the oracle's
DoF maps, boundary lists, vertices and transfer maps exactly."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "pmg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmg_[a-zA-Z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pmg):
    lib = ctypes.CDLL(pmg.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_gpu(pmg):
    lib = pmg.lib()
    assert lib.pmg_version() >= 10000
    # invalid arguments are rejected before any device call
    st = lib.pmg_box_dofmap(None, 2, 1, None)
    assert st == pmg.ERR_INVALID_ARG
    assert b"NULL" in lib.pmg_last_error()


@pytest.mark.parametrize("n,p,deform", [(2, 1, 0), (3, 2, 1), (2, 4, 0), (3, 3, 1)])
def test_box_helpers_match_oracle(pmg, orc, n, p, deform):
    b = pmg.box(n, n + 1, n + 2, deform=deform, amplitude=0.05)
    m = orc.mesh(n, n + 1, n + 2, deform=deform, amplitude=0.05)
    assert pmg.box_n_dofs(b, p) == orc.n_dofs(m, p)
    assert np.array_equal(pmg.box_dofmap(b, p, True), orc.dofmap(m, p, True))
    assert np.array_equal(pmg.box_dofmap(b, p, False), orc.dofmap(m, p, False))
    assert np.array_equal(pmg.box_boundary_dofs(b, p), orc.boundary_dofs(m, p))
    assert np.allclose(pmg.box_cell_vertices(b), orc.cell_vertices(m), rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 4])
def test_h_transfer_maps_match_oracle(pmg, orc, p):
    b, m = pmg.box(2, 3, 2), orc.mesh(2, 3, 2)
    mc, mf, w = pmg.box_h_transfer_maps(b, p)
    t = orc.Transfer(m, p, kind="h")
    assert np.array_equal(mc, t.map_c) and np.array_equal(mf, t.map_f)
    assert np.allclose(w, t.w, rtol=0, atol=1e-15)


@pytest.mark.parametrize("pc,pf", [(1, 2), (2, 4), (3, 6), (1, 3)])
def test_p_transfer_maps_match_oracle(pmg, orc, pc, pf):
    b, m = pmg.box(3), orc.mesh(3)
    mc, mf, w = pmg.box_p_transfer_maps(b, pc, pf)
    t = orc.Transfer(m, pc, pf, kind="p")
    assert np.array_equal(mc, t.map_c) and np.array_equal(mf, t.map_f)
    assert np.allclose(w, t.w, rtol=0, atol=1e-15)
