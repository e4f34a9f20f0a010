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


@pytest.mark.parametrize("cells", [(1, 1, 1), (2, 3, 2), (4, 1, 3)])
def test_box_dg_faces_cover_every_cell_face_once(pmg, cells):
    """Every (cell, face) pair appears exactly once; inner faces join
    neighbours along one axis with face+ = face- ^ 1 (standard orientation)."""
    b = pmg.box(*cells)
    inner, bdry = pmg.box_dg_faces(b)
    inner, bdry = inner.reshape(-1, 4), bdry.reshape(-1, 3)
    C = pmg.box_n_cells(b)
    seen = np.zeros((C, 6), dtype=int)
    np.add.at(seen, (inner[:, 0], inner[:, 2]), 1)
    np.add.at(seen, (inner[:, 1], inner[:, 3]), 1)
    np.add.at(seen, (bdry[:, 0], bdry[:, 1]), 1)
    assert np.all(seen == 1)
    assert np.all(inner[:, 3] == (inner[:, 2] ^ 1)) and np.all(inner[:, 2] % 2 == 1)
    nx, ny, _ = cells
    stride = np.array([1, nx, nx * ny])
    assert np.all(inner[:, 1] - inner[:, 0] == stride[inner[:, 2] // 2])
    assert np.all(bdry[:, 2] == 0)


def test_quadrature_points_match_oracle(pmg, orc):
    b = pmg.box(3, 2, 2, deform=1, amplitude=0.05)
    X = pmg.quadrature_points(b, 4)
    _, _, Xo = orc.geometry(orc.mesh(3, 2, 2, deform=1, amplitude=0.05), 4, want_x=True)
    assert np.abs(X - Xo.reshape(-1, 3)).max() < 1e-15


def test_cpp_header_compiles_and_links(pmg, tmp_path):
    """include/pmg.hpp (the C++ mirror) compiles standalone against the C ABI."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no g++")
    src = tmp_path / "use.cpp"
    src.write_text('#include "pmg.hpp"\n'
                   'int main() { pmg_box b{}; b.nx = b.ny = b.nz = 2; b.hi[0] = b.hi[1] = b.hi[2] = 1;\n'
                   '  return pmg_box_n_cells(&b) == 8 ? 0 : 1; }\n')
    exe = tmp_path / "use"
    pkg = os.path.dirname(pmg.LIB_PATH)
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", pkg, "-lpmg", f"-Wl,-rpath,{pkg}"],
                   check=True)
    assert subprocess.run([str(exe)]).returncode == 0
