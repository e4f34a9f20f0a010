"""MI355X-native matrix-free multigrid hot path (Python host mirror).

Thin ctypes layer over libpmg.so (gfx950 HIP kernels + C++ host classes,
C ABI in include/pmg.h).  The classes mirror the reference's interfaces
(gshubhamk/portable-multigrid):

  LaplaceOperatorBK3   include/operators/portable_laplace_operator_bk3.h
  LaplaceOperatorDG    include/operators/portable_laplace_operator_dg.h
  GeometricTransfer    include/multigrid/portable_geometric_transfer.h
  PolynomialTransfer   include/multigrid/portable_polynomial_transfer.h
  ContinuousTransfer   include/multigrid/portable_continuous_transfer.h
  PreconditionChebyshev  deal.II PreconditionChebyshev as configured in
                       tests/poisson_cube_bk3/program.cc:474-515
  VCycleMultigrid      include/multigrid/portable_v_cycle_multigrid.h
  solve_cg             tests/common/portable_multigrid_solver.h:147-169

Vectors are torch tensors on the GPU (float64 for FP64 handles, float32 for
FP32 handles); torch is only the device-memory and stream provider.  There is
no CPU fallback: if libpmg.so is missing every call raises.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmg.so")

FP64, FP32 = 0, 1
TRANSFER_H, TRANSFER_P, TRANSFER_DG_CG = 0, 1, 2
PRECOND_NONE, PRECOND_JACOBI, PRECOND_MG = 0, 1, 2
COARSE_CHEBYSHEV, COARSE_PCG = 0, 1
SCHEDULE_ATOMIC, SCHEDULE_COLORED, SCHEDULE_PATCH, SCHEDULE_GATHER, SCHEDULE_GRID = 0, 1, 2, 3, 4
TRANSFER_LAYOUT_WEIGHTED, TRANSFER_LAYOUT_OWNER_MAP, TRANSFER_LAYOUT_OWNER_LINES = 0, 1, 2
INVALID_DOF = 0xFFFFFFFF
# index arrays of pmg_operator_index_data / pmg_transfer_index_data: name -> (id, dtype)
OPERATOR_INDEX = {"map": (0, np.uint32), "constrained": (1, np.uint32),
                  "cell_order": (2, np.uint32), "lines": (3, np.uint32),
                  "line_flags": (4, np.uint8), "patch_dofs": (5, np.uint32),
                  "patch_counts": (6, np.uint32), "patch_slots": (7, np.uint16),
                  "patch_lines": (8, np.uint16), "patch_line_flags": (9, np.uint8),
                  "vertices": (10, np.float64), "patch_colors": (11, np.uint8),
                  "patch_ncolors": (12, np.uint8), "patch_chunks": (13, np.uint32)}
TRANSFER_INDEX = {"coarse_map": (0, np.uint32), "fine_map": (1, np.uint32),
                  "fine_lines": (2, np.uint32), "restrict_offsets": (3, np.uint32),
                  "restrict_positions": (4, np.uint32), "dg_offsets": (5, np.uint32),
                  "dg_list": (6, np.uint32)}

OK, ERR_INVALID_ARG, ERR_SIZE, ERR_HIP, ERR_NOT_CONVERGED, ERR_UNSUPPORTED, ERR_ALLOC = range(7)

_lib = None


class PMGError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libpmg status {status}: {msg}")
        self.status = status


def build(force=False, jobs=8):
    """Compile libpmg.so for gfx950 with the package Makefile (hipcc)."""
    cmd = ["make", "-s", "-C", _HERE, f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return LIB_PATH


class Box(ctypes.Structure):
    """Structured box mesh description (pmg_box)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("deform", ctypes.c_int), ("amplitude", ctypes.c_double)]


def box(nx, ny=None, nz=None, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0, amplitude=0.0):
    b = Box()
    b.nx, b.ny, b.nz = nx, nx if ny is None else ny, nx if nz is None else nz
    b.lo[:] = lo
    b.hi[:] = hi
    b.deform, b.amplitude = deform, amplitude
    return b


def refine(b):
    return box(2 * b.nx, 2 * b.ny, 2 * b.nz, tuple(b.lo), tuple(b.hi), b.deform, b.amplitude)


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double
_BP = ctypes.POINTER(Box)


def lib():
    """Load libpmg.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run portable-multigrid_amd build() "
                           "(hipcc --offload-arch=gfx950) first; there is no CPU fallback")
    # torch brings its own libamdhip64.so.7; loading it first makes libpmg.so
    # bind to that same HIP runtime, so the library and torch share devices,
    # streams and allocations (a second runtime loaded ahead of torch's leaves
    # torch without visible GPUs)
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        "pmg_last_error": (ctypes.c_char_p, []),
        "pmg_version": (_i, []),
        "pmg_device_synchronize": (_i, []),
        "pmg_box_n_cells": (_i64, [_BP]),
        "pmg_box_n_dofs": (_i64, [_BP, _i]),
        "pmg_box_cell_vertices": (_i, [_BP, _vp]),
        "pmg_box_dofmap": (_i, [_BP, _i, _i, _vp]),
        "pmg_box_boundary_dofs": (_i, [_BP, _i, _vp, P(_i64)]),
        "pmg_box_h_transfer_maps": (_i, [_BP, _i, _vp, _vp, _vp]),
        "pmg_box_p_transfer_maps": (_i, [_BP, _i, _i, _vp, _vp, _vp]),
        "pmg_box_dg_faces": (_i, [_BP, _vp, P(_i64), _vp, P(_i64)]),
        "pmg_dg_operator_create": (_i, [_i, _i, _i, _i, _i64, _vp, _i64, _vp, _i64, _vp, _vp,
                                        _i, _i, P(_vp)]),
        "pmg_dg_operator_penalty": (_i, [_vp, _vp]),
        "pmg_dg_operator_create_box": (_i, [_BP, _i, _i, _i, _i, _i, _i, P(_vp)]),
        "pmg_operator_create": (_i, [_i, _i, _i, _i, _i64, _vp, _i64, _vp, _i64, _vp, P(_vp)]),
        "pmg_operator_create_from_G": (_i, [_i, _i, _i, _i, _i64, _vp, _i64, _vp, _i64, _vp,
                                            P(_vp)]),
        "pmg_operator_create_box": (_i, [_BP, _i, _i, _i, P(_vp)]),
        "pmg_operator_index_data": (_i, [_vp, _i, _vp, P(_i64)]),
        "pmg_transfer_create_box": (_i, [_i, _i, _BP, _i, _i, _d, P(_vp)]),
        "pmg_transfer_index_data": (_i, [_vp, _i, _vp, P(_i64)]),
        "pmg_operator_destroy": (_i, [_vp]),
        "pmg_operator_m": (_i, [_vp, P(_i64)]),
        "pmg_operator_info": (_i, [_vp, P(_i), P(_i), P(_i), P(_i64)]),
        "pmg_operator_vmult": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_operator_Tvmult": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_operator_apply_add": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_operator_set_subdomain": (_i, [_vp, _vp, _vp, _i64, _vp, _i64]),
        "pmg_transfer_create_from_constraints": (_i, [_i, _i, _i, _i, _i64, _vp, _vp, _vp, _i64,
                                                      _i64, _i64, _d, P(_vp)]),
        "pmg_operator_vmult_neumann": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_operator_vmult_interface_cells": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_operator_compute_diagonal": (_i, [_vp]),
        "pmg_operator_diagonal_inverse": (_i, [_vp, _vp, _vp]),
        "pmg_operator_diagonal_entry": (_i, [_vp, _i64, P(_d)]),
        "pmg_operator_vmult_dummy": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
        "pmg_transfer_reinit": (_i, [_vp, _i64, _vp, _vp, _vp, _i64, _i64, _i64]),
        "pmg_cg_solve_noalloc": (_i, [_vp, _i, _vp, _vp, _vp, _d, _i, P(_i), P(_d), P(_d),
                                      _vp]),
        "pmg_operator_set_coefficient": (_i, [_vp, _vp, _vp]),
        "pmg_operator_use_affine_geometry": (_i, [_vp, _i]),
        "pmg_operator_gradient_coefficient": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "pmg_operator_algorithmic_bytes": (_i, [_vp, P(_d), P(_d)]),
        "pmg_operator_assemble_rhs": (_i, [_vp, _i, _i, _vp, _vp]),
        "pmg_operator_set_schedule": (_i, [_vp, _i, _i]),
        "pmg_operator_set_box_structure": (_i, [_vp, _i, _i, _i]),
        "pmg_operator_l2_error": (_i, [_vp, _vp, _i, _i, P(_d), _vp]),
        "pmg_transfer_create": (_i, [_i, _i, _i, _i, _i64, _vp, _vp, _vp, _i64, _i64, _d,
                                     P(_vp)]),
        "pmg_transfer_destroy": (_i, [_vp]),
        "pmg_transfer_info": (_i, [_vp, P(_i), P(_i), P(_i), P(_i), P(_i)]),
        "pmg_transfer_prolongate_and_add": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_transfer_restrict_and_add": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_chebyshev_create": (_i, [_vp, _i, _d, _i, P(_vp)]),
        "pmg_chebyshev_destroy": (_i, [_vp]),
        "pmg_chebyshev_vmult": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_chebyshev_step": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_chebyshev_info": (_i, [_vp, P(_d), P(_d), P(_i)]),
        "pmg_multigrid_create": (_i, [_i, _vp, _vp, _vp, _i, _d, _i, P(_vp)]),
        "pmg_multigrid_destroy": (_i, [_vp]),
        "pmg_multigrid_vmult": (_i, [_vp, _vp, _vp, _vp]),
        "pmg_multigrid_uses_graph": (_i, [_vp, P(_i)]),
        "pmg_multigrid_level_times": (_i, [_vp, _vp, _i, _vp, _vp]),
        "pmg_cg_solve": (_i, [_vp, _i, _vp, _vp, _vp, _d, _i, P(_i), P(_d), P(_d), _vp]),
        "pmg_vector_dot": (_i, [_i, _i64, _vp, _vp, P(_d), _vp]),
        "pmg_vector_convert": (_i, [_i64, _i, _vp, _i, _vp, _vp]),
        "pmg_measure_fma_rate": (_i, [_i, P(_d), _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(status):
    if status != OK:
        raise PMGError(status, lib().pmg_last_error().decode())


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_ptr(t, precision, n=None):
    import torch
    want = torch.float64 if precision == FP64 else torch.float32
    if not (t.is_cuda and t.dtype == want and t.is_contiguous()):
        raise PMGError(ERR_INVALID_ARG, f"expected a contiguous CUDA {want} tensor, got "
                       f"{t.dtype} on {t.device}")
    if n is not None and t.numel() != n:
        raise PMGError(ERR_SIZE, f"vector has {t.numel()} entries, expected {n}")
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ meshes
def box_n_cells(b):
    return lib().pmg_box_n_cells(ctypes.byref(b))


def box_n_dofs(b, p):
    return lib().pmg_box_n_dofs(ctypes.byref(b), p)


def box_cell_vertices(b):
    out = np.zeros(box_n_cells(b) * 24)
    _check(lib().pmg_box_cell_vertices(ctypes.byref(b), _np_ptr(out)))
    return out


def box_dofmap(b, p, dirichlet=True):
    out = np.zeros(box_n_cells(b) * (p + 1) ** 3, dtype=np.uint32)
    _check(lib().pmg_box_dofmap(ctypes.byref(b), p, int(dirichlet), _np_ptr(out)))
    return out


def box_boundary_dofs(b, p):
    n = ctypes.c_int64()
    _check(lib().pmg_box_boundary_dofs(ctypes.byref(b), p, None, ctypes.byref(n)))
    out = np.zeros(n.value, dtype=np.uint32)
    _check(lib().pmg_box_boundary_dofs(ctypes.byref(b), p, _np_ptr(out), ctypes.byref(n)))
    return out


def box_h_transfer_maps(coarse, p, weights=True):
    """Coarse map, fine patch map and (weights=True, else None) the host
    weights 1/multiplicity of the h-transfer coarse -> refine(coarse)."""
    C = box_n_cells(coarse)
    mc = np.zeros(C * (p + 1) ** 3, dtype=np.uint32)
    mf = np.zeros(C * (2 * p + 1) ** 3, dtype=np.uint32)
    w = np.zeros(C * (2 * p + 1) ** 3) if weights else None
    _check(lib().pmg_box_h_transfer_maps(ctypes.byref(coarse), p, _np_ptr(mc), _np_ptr(mf),
                                         None if w is None else _np_ptr(w)))
    return mc, mf, w


def box_p_transfer_maps(b, pc, pf, weights=True):
    C = box_n_cells(b)
    mc = np.zeros(C * (pc + 1) ** 3, dtype=np.uint32)
    mf = np.zeros(C * (pf + 1) ** 3, dtype=np.uint32)
    w = np.zeros(C * (pf + 1) ** 3) if weights else None
    _check(lib().pmg_box_p_transfer_maps(ctypes.byref(b), pc, pf, _np_ptr(mc), _np_ptr(mf),
                                         None if w is None else _np_ptr(w)))
    return mc, mf, w


def box_dg_faces(b):
    """DG face lists of a box: inner [F][4] = (cell-, cell+, face-, face+),
    boundary [Fb][3] = (cell, face, boundary id 0)."""
    ni, nb = ctypes.c_int64(), ctypes.c_int64()
    _check(lib().pmg_box_dg_faces(ctypes.byref(b), None, ctypes.byref(ni), None, ctypes.byref(nb)))
    inner = np.zeros(4 * ni.value, dtype=np.uint32)
    bdry = np.zeros(3 * nb.value, dtype=np.uint32)
    _check(lib().pmg_box_dg_faces(ctypes.byref(b), _np_ptr(inner), ctypes.byref(ni),
                                  _np_ptr(bdry), ctypes.byref(nb)))
    return inner, bdry


def _index_data(fn, h, spec):
    which, dtype = spec
    n = ctypes.c_int64(0)
    _check(fn(h, which, None, ctypes.byref(n)))
    out = np.empty(n.value, dtype=dtype)
    if n.value:
        _check(fn(h, which, _np_ptr(out), ctypes.byref(n)))
    return out


# ------------------------------------------------------------------ operator
class LaplaceOperatorBK3:
    """Continuous Q_p Laplace operator on the GPU (LaplaceOperatorBK3<3,p,number>).

    Build from a structured box (``LaplaceOperatorBK3(box, p)``) or from raw
    arrays (``LaplaceOperatorBK3.from_arrays``)."""

    def __init__(self, b=None, degree=1, n_q_points_1d=None, precision=FP64, _raw=None):
        self.degree = degree
        self.nq = degree + 1 if n_q_points_1d is None else n_q_points_1d
        self.precision = precision
        self.h = ctypes.c_void_p()
        if _raw is None:
            # the box's map, constraints and vertices are generated on the device;
            # a requested non-default cell order (PMG_CELL_ORDER) goes through
            # the host arrays
            self.box = b
            n_cells, n_dofs = box_n_cells(b), box_n_dofs(b, degree)
            if os.environ.get("PMG_CELL_ORDER", "caller") == "caller":
                _check(lib().pmg_operator_create_box(ctypes.byref(b), degree, self.nq, precision,
                                                     ctypes.byref(self.h)))
            else:
                dofmap = box_dofmap(b, degree, True)
                cons = box_boundary_dofs(b, degree)
                cv = box_cell_vertices(b)
                _check(lib().pmg_operator_create(3, degree, self.nq, precision, n_cells,
                                                 _np_ptr(dofmap), n_dofs, _np_ptr(cons),
                                                 len(cons), _np_ptr(cv), ctypes.byref(self.h)))
        else:
            dofmap, n_dofs, cons, cv, G = _raw
            dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
            cons = np.ascontiguousarray(cons, dtype=np.uint32)
            n_cells = len(dofmap) // (degree + 1) ** 3
            if G is not None:
                G = np.ascontiguousarray(G, dtype=np.float64)
                _check(lib().pmg_operator_create_from_G(3, degree, self.nq, precision, n_cells,
                                                        _np_ptr(dofmap), n_dofs, _np_ptr(cons),
                                                        len(cons), _np_ptr(G),
                                                        ctypes.byref(self.h)))
            else:
                cv = np.ascontiguousarray(cv, dtype=np.float64)
                _check(lib().pmg_operator_create(3, degree, self.nq, precision, n_cells,
                                                 _np_ptr(dofmap), n_dofs, _np_ptr(cons),
                                                 len(cons), _np_ptr(cv), ctypes.byref(self.h)))
        self.n_cells = n_cells
        self._m = n_dofs

    @classmethod
    def from_arrays(cls, degree, dofmap, n_dofs, constrained, cell_vertices=None, G=None,
                    n_q_points_1d=None, precision=FP64):
        return cls(None, degree, n_q_points_1d, precision,
                   _raw=(dofmap, n_dofs, constrained, cell_vertices, G))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.pmg_operator_destroy(self.h)
            self.h = None

    def m(self):
        return self._m

    def n(self):
        return self._m

    @property
    def dtype(self):
        import torch
        return torch.float64 if self.precision == FP64 else torch.float32

    def initialize_dof_vector(self):
        import torch
        return torch.zeros(self._m, dtype=self.dtype, device="cuda")

    def vmult(self, dst, src):
        _check(lib().pmg_operator_vmult(self.h, _dev_ptr(dst, self.precision, self._m),
                                        _dev_ptr(src, self.precision, self._m), _stream()))

    def Tvmult(self, dst, src):
        _check(lib().pmg_operator_Tvmult(self.h, _dev_ptr(dst, self.precision, self._m),
                                         _dev_ptr(src, self.precision, self._m), _stream()))

    def set_subdomain(self, dofmap, physical_dofs, interface_cells):
        """SubdomainLaplaceOperator index sets: the plain map (caller cell
        order) with only the physical boundary DoFs eliminated, and the
        interface cells (caller indices)."""
        dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
        phys = np.ascontiguousarray(physical_dofs, dtype=np.uint32)
        cells = np.ascontiguousarray(interface_cells, dtype=np.uint32)
        _check(lib().pmg_operator_set_subdomain(self.h, _np_ptr(dofmap), _np_ptr(phys), len(phys),
                                                _np_ptr(cells), len(cells)))

    def vmult_neumann(self, dst, src):
        _check(lib().pmg_operator_vmult_neumann(self.h, _dev_ptr(dst, self.precision, self._m),
                                                _dev_ptr(src, self.precision, self._m), _stream()))

    def vmult_interface_cell_range(self, dst, src):
        _check(lib().pmg_operator_vmult_interface_cells(
            self.h, _dev_ptr(dst, self.precision, self._m), _dev_ptr(src, self.precision, self._m),
            _stream()))

    def apply_add(self, dst, src):
        _check(lib().pmg_operator_apply_add(self.h, _dev_ptr(dst, self.precision, self._m),
                                            _dev_ptr(src, self.precision, self._m), _stream()))

    def compute_diagonal(self):
        _check(lib().pmg_operator_compute_diagonal(self.h))

    def el(self, row, col):
        """Diagonal entry A_ii = 1 / (inverse diagonal)_ii (row == col only)."""
        if row != col:
            raise PMGError(ERR_UNSUPPORTED, "el: only diagonal entries")
        v = ctypes.c_double()
        _check(lib().pmg_operator_diagonal_entry(self.h, int(row), ctypes.byref(v)))
        return v.value

    def vmult_dummy(self, dst, src, ghost_exchange_on, computation_on):
        _check(lib().pmg_operator_vmult_dummy(self.h, _dev_ptr(dst, self.precision, self._m),
                                              _dev_ptr(src, self.precision, self._m),
                                              int(bool(ghost_exchange_on)),
                                              int(bool(computation_on)), _stream()))

    def get_matrix_diagonal_inverse(self):
        out = self.initialize_dof_vector()
        _check(lib().pmg_operator_diagonal_inverse(self.h, _dev_ptr(out, self.precision),
                                                   _stream()))
        return out

    def set_coefficient(self, kappa):
        """kappa: CUDA float64 tensor [n_cells * nq^3] or None (kappa = 1)."""
        if kappa is None:
            _check(lib().pmg_operator_set_coefficient(self.h, None, _stream()))
        else:
            _check(lib().pmg_operator_set_coefficient(
                self.h, _dev_ptr(kappa, FP64, self.n_cells * self.nq ** 3), _stream()))

    def use_affine_geometry(self, enable=True):
        """vmult from 6 values per affine cell instead of G at every q-point
        (raises PMGError if a cell is not affine or a coefficient is set)."""
        _check(lib().pmg_operator_use_affine_geometry(self.h, 1 if enable else 0))

    def gradient_coefficient(self, u, lam, out):
        _check(lib().pmg_operator_gradient_coefficient(
            self.h, _dev_ptr(u, self.precision, self._m), _dev_ptr(lam, self.precision, self._m),
            _dev_ptr(out, self.precision, self.n_cells * self.nq ** 3), _stream()))

    def index_data(self, name):
        """Copy of the device index array `name` (OPERATOR_INDEX) as numpy; an
        empty array where it is not built (patch arrays: once SCHEDULE_PATCH
        was selected)."""
        return _index_data(lib().pmg_operator_index_data, self.h, OPERATOR_INDEX[name])

    def set_box_structure(self, nx, ny, nz):
        """Declare the cells a lexicographic nx x ny x nz box (verified on the
        device); enables SCHEDULE_GRID."""
        _check(lib().pmg_operator_set_box_structure(self.h, int(nx), int(ny), int(nz)))

    def set_schedule(self, schedule=SCHEDULE_ATOMIC, g_load=-1):
        """Cell-loop schedule: SCHEDULE_ATOMIC (per-cell device atomics),
        SCHEDULE_COLORED (one launch per colour) or SCHEDULE_PATCH (batch
        patches in LDS).  The constructor's default is PATCH for
        n_q_points_1d <= 5 (FP64) or <= 7 (FP32), ATOMIC otherwise.  g_load:
        G load policy (0 early non-temporal, 1 early, 2 late non-temporal,
        -1 the schedule's default)."""
        _check(lib().pmg_operator_set_schedule(self.h, int(schedule), g_load))

    def assemble_rhs(self, f_kind=1, n_q_points_1d=0):
        """b = int phi_i f (f_kind 0: f = 1, 1: 3 pi^2 sin sin sin), constrained rows 0."""
        b = self.initialize_dof_vector()
        _check(lib().pmg_operator_assemble_rhs(self.h, f_kind, n_q_points_1d,
                                               _dev_ptr(b, self.precision), _stream()))
        return b

    def l2_error(self, u, u_kind=1, n_q_points_1d=0):
        e = ctypes.c_double()
        _check(lib().pmg_operator_l2_error(self.h, _dev_ptr(u, self.precision, self._m), u_kind,
                                           n_q_points_1d, ctypes.byref(e), _stream()))
        return e.value

    def algorithmic_bytes(self):
        b, f = ctypes.c_double(), ctypes.c_double()
        _check(lib().pmg_operator_algorithmic_bytes(self.h, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value


class LaplaceOperatorDG(LaplaceOperatorBK3):
    """Symmetric interior penalty DG Laplacian on the GPU (LaplaceOperatorDG<3,p,nq,number>).

    Cell-major DoFs.  penalty: None (computed from penalty_kind, 0 = the
    reference's p(p+1)/2 (1/h- + 1/h+), 1 = (p+1)^2 (1/h- + 1/h+)) or an
    array over the faces (inner faces first).  metric 0 uses J^{-1}n, 1 the
    J^{-T}n of the reference's GPU face data.  Boundary faces of a box are
    weak homogeneous Dirichlet unless neumann_sides (bit f = box side f) says
    otherwise."""

    def __init__(self, b=None, degree=1, n_q_points_1d=None, precision=FP64, penalty=None,
                 penalty_kind=0, metric=0, neumann_sides=0, _raw=None, device_setup=True):
        """device_setup (a box with the computed penalty): vertices, faces,
        penalty and geometry built on the device (pmg_dg_operator_create_box);
        False builds the face lists on the host and passes them in."""
        self.degree = degree
        self.nq = degree + 1 if n_q_points_1d is None else n_q_points_1d
        self.precision = precision
        self.h = ctypes.c_void_p()
        if _raw is None and penalty is None and device_setup:
            self.box = b
            _check(lib().pmg_dg_operator_create_box(ctypes.byref(b), degree, self.nq, precision,
                                                    penalty_kind, metric, neumann_sides,
                                                    ctypes.byref(self.h)))
            self.n_cells = box_n_cells(b)
            n = (b.nx, b.ny, b.nz)
            self.n_faces = (3 * n[0] * n[1] * n[2] - n[1] * n[2] - n[0] * n[2] - n[0] * n[1]
                            + 2 * (n[1] * n[2] + n[0] * n[2] + n[0] * n[1]))
            self._m = self.n_cells * (degree + 1) ** 3
            return
        if _raw is None:
            self.box = b
            cv = box_cell_vertices(b)
            inner, bdry = box_dg_faces(b)
            if neumann_sides:
                f = bdry.reshape(-1, 3)
                f[:, 2] = ((neumann_sides >> f[:, 1].astype(np.int64)) & 1).astype(np.uint32)
        else:
            cv, inner, bdry = _raw
            cv = np.ascontiguousarray(cv, dtype=np.float64)
            inner = np.ascontiguousarray(inner, dtype=np.uint32)
            bdry = np.ascontiguousarray(bdry, dtype=np.uint32)
        self.n_cells = len(cv) // 24
        self.n_faces = len(inner) // 4 + len(bdry) // 3
        pen = None if penalty is None else np.ascontiguousarray(penalty, dtype=np.float64)
        _check(lib().pmg_dg_operator_create(3, degree, self.nq, precision, self.n_cells,
                                            _np_ptr(cv), len(inner) // 4, _np_ptr(inner),
                                            len(bdry) // 3, _np_ptr(bdry),
                                            None if pen is None else _np_ptr(pen), penalty_kind,
                                            metric, ctypes.byref(self.h)))
        self._m = self.n_cells * (degree + 1) ** 3

    @classmethod
    def from_arrays(cls, degree, cell_vertices, inner_faces, boundary_faces, n_q_points_1d=None,
                    precision=FP64, penalty=None, penalty_kind=0, metric=0):
        return cls(None, degree, n_q_points_1d, precision, penalty, penalty_kind, metric,
                   _raw=(cell_vertices, inner_faces, boundary_faces))

    def penalty(self):
        out = np.zeros(self.n_faces)
        _check(lib().pmg_dg_operator_penalty(self.h, _np_ptr(out)))
        return out


# ------------------------------------------------------------------ transfers
class _Transfer:
    def __init__(self, kind, precision, pc, pf, mc, mf, w, n_c, n_f, scale, cons=None, box=None):
        """cons (with w None): the fine constrained DoFs, weights built on the device.
        box (maps None): the (coarse) box, maps and weights generated on the device."""
        self.precision = precision
        self.n_c, self.n_f = n_c, n_f
        self.h = ctypes.c_void_p()
        if box is not None:
            _check(lib().pmg_transfer_create_box(kind, precision, ctypes.byref(box), pc, pf,
                                                 scale, ctypes.byref(self.h)))
            return
        mc = np.ascontiguousarray(mc, dtype=np.uint32)
        mf = None if mf is None else np.ascontiguousarray(mf, dtype=np.uint32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        n_cells = len(mc) // (pc + 1) ** 3
        if w is None and cons is not None:
            cons = np.ascontiguousarray(cons, dtype=np.uint32)
            _check(lib().pmg_transfer_create_from_constraints(
                kind, precision, pc, pf, n_cells, _np_ptr(mc), None if mf is None else _np_ptr(mf),
                _np_ptr(cons), len(cons),
                n_c, n_f, scale, ctypes.byref(self.h)))
            return
        _check(lib().pmg_transfer_create(kind, precision, pc, pf, n_cells, _np_ptr(mc),
                                         None if mf is None else _np_ptr(mf),
                                         None if w is None else _np_ptr(w), n_c, n_f, scale,
                                         ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.pmg_transfer_destroy(self.h)
            self.h = None

    def prolongate_and_add(self, dst, src):
        _check(lib().pmg_transfer_prolongate_and_add(
            self.h, _dev_ptr(dst, self.precision, self.n_f), _dev_ptr(src, self.precision, self.n_c),
            _stream()))

    def restrict_and_add(self, dst, src):
        _check(lib().pmg_transfer_restrict_and_add(
            self.h, _dev_ptr(dst, self.precision, self.n_c), _dev_ptr(src, self.precision, self.n_f),
            _stream()))

    def reinit(self, mc, mf, cons, n_c, n_f):
        """MGTransferBase::reinit: new maps, fine constraints (device weights) and sizes."""
        mc = np.ascontiguousarray(mc, dtype=np.uint32)
        mf = None if mf is None else np.ascontiguousarray(mf, dtype=np.uint32)
        cons = np.ascontiguousarray(cons if cons is not None else [], dtype=np.uint32)
        pc = self.info()[2]
        n_cells = len(mc) // (pc + 1) ** 3
        _check(lib().pmg_transfer_reinit(self.h, n_cells, _np_ptr(mc),
                                         None if mf is None else _np_ptr(mf), _np_ptr(cons),
                                         len(cons), n_c, n_f))
        self.n_c, self.n_f = n_c, n_f

    def index_data(self, name):
        """Copy of the device index array `name` (TRANSFER_INDEX) as numpy."""
        return _index_data(lib().pmg_transfer_index_data, self.h, TRANSFER_INDEX[name])

    def info(self):
        """(kind, precision, degree_coarse, degree_fine, layout)"""
        v = [ctypes.c_int() for _ in range(5)]
        _check(lib().pmg_transfer_info(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def layout(self):
        """TRANSFER_LAYOUT_WEIGHTED, _OWNER_MAP or _OWNER_LINES (fine map storage)."""
        return self.info()[4]


class GeometricTransfer(_Transfer):
    """h-transfer coarse box -> once-refined box at degree p (GeometricTransfer<3,p>).

    prolongation_scale = 2.0 reproduces the reference HEAD's doubled
    h-prolongation (portable_geometric_transfer.h:805-855)."""

    def __init__(self, coarse_box, degree, precision=FP64, prolongation_scale=1.0,
                 device_weights=False):
        """device_weights: maps and weights generated on the device
        (pmg_transfer_create_box); else host maps with explicit weights."""
        fine = refine(coarse_box)
        n_c, n_f = box_n_dofs(coarse_box, degree), box_n_dofs(fine, degree)
        if device_weights:
            super().__init__(TRANSFER_H, precision, degree, degree, None, None, None, n_c, n_f,
                             prolongation_scale, box=coarse_box)
            return
        mc, mf, w = box_h_transfer_maps(coarse_box, degree)
        super().__init__(TRANSFER_H, precision, degree, degree, mc, mf, w, n_c, n_f,
                         prolongation_scale)


class PolynomialTransfer(_Transfer):
    """p-transfer on one box, degree pc -> pf (PolynomialTransfer<3,pc,pf>)."""

    def __init__(self, b, pc, pf, precision=FP64, device_weights=False):
        n_c, n_f = box_n_dofs(b, pc), box_n_dofs(b, pf)
        if device_weights:
            super().__init__(TRANSFER_P, precision, pc, pf, None, None, None, n_c, n_f, 1.0, box=b)
            return
        mc, mf, w = box_p_transfer_maps(b, pc, pf)
        super().__init__(TRANSFER_P, precision, pc, pf, mc, mf, w, n_c, n_f, 1.0)


class ContinuousTransfer(_Transfer):
    """CG(p) -> DG(p) on one box (ContinuousTransfer<3,p>): the DG DoFs of a
    cell take the values of the CG DoFs at the same support points; the
    restriction sums them back (constrained CG DoFs skipped)."""

    def __init__(self, b, p, precision=FP64):
        super().__init__(TRANSFER_DG_CG, precision, p, p, None, None, None, box_n_dofs(b, p),
                         box_n_cells(b) * (p + 1) ** 3, 1.0, box=b)


# ------------------------------------------------------------------ smoother
class PreconditionChebyshev:
    """Chebyshev smoother over the operator's inverse diagonal (deal.II semantics)."""

    def __init__(self, op, degree=3, smoothing_range=15.0, eig_cg_n_iterations=10):
        self.op = op
        self.h = ctypes.c_void_p()
        _check(lib().pmg_chebyshev_create(op.h, degree, smoothing_range, eig_cg_n_iterations,
                                          ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.pmg_chebyshev_destroy(self.h)
            self.h = None

    def vmult(self, dst, src):
        p, m = self.op.precision, self.op.m()
        _check(lib().pmg_chebyshev_vmult(self.h, _dev_ptr(dst, p, m), _dev_ptr(src, p, m),
                                         _stream()))

    def step(self, dst, src):
        p, m = self.op.precision, self.op.m()
        _check(lib().pmg_chebyshev_step(self.h, _dev_ptr(dst, p, m), _dev_ptr(src, p, m),
                                        _stream()))

    def info(self):
        mx, mn, deg = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(lib().pmg_chebyshev_info(self.h, ctypes.byref(mx), ctypes.byref(mn),
                                        ctypes.byref(deg)))
        return mx.value, mn.value, deg.value


# ------------------------------------------------------------------ multigrid
class VCycleMultigrid:
    """V-cycle over levels 0 (coarsest) .. L-1 (VCycleMultigrid / MultigridSolver::vmult)."""

    def __init__(self, operators, transfers, smoothers, coarse=COARSE_PCG, coarse_tol=1e-12,
                 outer_precision=FP64):
        L = len(operators)
        self.L = L
        self._keep = (operators, transfers, smoothers)
        self.outer_precision = outer_precision
        self.n_fine = operators[-1].m()
        ops = (ctypes.c_void_p * L)(*[o.h.value for o in operators])
        trs = (ctypes.c_void_p * L)(*[(t.h.value if t is not None else None) for t in transfers])
        sms = (ctypes.c_void_p * L)(*[(s.h.value if s is not None else None) for s in smoothers])
        self.h = ctypes.c_void_p()
        _check(lib().pmg_multigrid_create(L, ops, trs, sms, coarse, coarse_tol, outer_precision,
                                          ctypes.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.pmg_multigrid_destroy(self.h)
            self.h = None

    def vmult(self, dst, src):
        p = self.outer_precision
        _check(lib().pmg_multigrid_vmult(self.h, _dev_ptr(dst, p, self.n_fine),
                                         _dev_ptr(src, p, self.n_fine), _stream()))

    LEVEL_STAGES = ("pre_smooth", "residual_matvec", "restrict", "coarse_solve", "prolongate",
                    "post_smooth")

    def level_times(self, src, n_applications=5):
        """Seconds per application of each stage at each level (eager V-cycles,
        device events): array [levels][6] in the order of LEVEL_STAGES."""
        out = np.zeros(self.L * 6)
        _check(lib().pmg_multigrid_level_times(self.h, _dev_ptr(src, self.outer_precision,
                                                                 self.n_fine),
                                               int(n_applications), _np_ptr(out), _stream()))
        return out.reshape(self.L, 6)

    def uses_graph(self):
        """True if the V-cycle body replays a captured HIP graph."""
        v = ctypes.c_int()
        _check(lib().pmg_multigrid_uses_graph(self.h, ctypes.byref(v)))
        return bool(v.value)


def solve_cg(op, b, x, precond=PRECOND_JACOBI, mg=None, rel_tol=1e-10, max_it=1000,
             raise_on_failure=True):
    """Preconditioned CG (deal.II SolverCG semantics); returns (iterations, res0, res)."""
    its, r0, r = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    st = lib().pmg_cg_solve(op.h, precond, mg.h if mg is not None else None,
                            _dev_ptr(b, op.precision, op.m()), _dev_ptr(x, op.precision, op.m()),
                            rel_tol, max_it, ctypes.byref(its), ctypes.byref(r0), ctypes.byref(r),
                            _stream())
    if st != OK and (raise_on_failure or st != ERR_NOT_CONVERGED):
        _check(st)
    return its.value, r0.value, r.value


def measure_fma_rate(precision=FP64):
    """Device vector FMA rate in flop/s (measurement only)."""
    v = ctypes.c_double()
    _check(lib().pmg_measure_fma_rate(precision, ctypes.byref(v), _stream()))
    return v.value


def vector_dot(a, b, precision=FP64):
    out = ctypes.c_double()
    _check(lib().pmg_vector_dot(precision, a.numel(), _dev_ptr(a, precision),
                                _dev_ptr(b, precision), ctypes.byref(out), _stream()))
    return out.value


# ------------------------------------------------------------------ hierarchies
def build_hierarchy(levels, n_q_points_1d=None, precision=FP64, smoother_degree=3,
                    smoothing_range=15.0, eig_cg_n_iterations=10, coarse=COARSE_PCG,
                    coarse_tol=1e-12, prolongation_scale=1.0, outer_precision=None,
                    dg_penalty_kind=0, dg_metric=0, coefficient=None, coarse_precision=None,
                    affine=False):
    """V-cycle over `levels` = [(box, degree) or (box, degree, "dg"), ...]
    from coarsest to finest.

    Consecutive levels on the same box are joined by a PolynomialTransfer
    (degree change) or a ContinuousTransfer (same degree, CG -> DG), levels
    whose box is refined once by a GeometricTransfer (same degree), as the
    reference's ph hierarchies (tests/poisson_cube_bk3/program.cc:281-351,
    421-470; tests/poisson_cube_bk3_fp64_dg/program.cc:294-460).
    n_q_points_1d: None -> degree+1 per level, or a callable degree -> nq.
    coefficient: None, or a callable (box, n_q_points_1d) -> CUDA float64
    tensor kappa [n_cells * nq^3] applied to each continuous level before its
    diagonal and smoother are set up.
    coarse_precision: the coarsest operator's precision.  None: FP64 when FP32
    levels precondition an FP64 outer solve (the coarse-grid solve stays in
    FP64, like the diagonal inversion and the outer CG), else `precision`.
    affine: the continuous levels' vmults use the affine geometry (6 values
    per cell; every cell must be affine, and no coefficient)."""
    if coarse_precision is None:
        mixed = precision == FP32 and outer_precision == FP64 and len(levels) > 1
        coarse_precision = FP64 if mixed else precision
    ops, trs, sms = [], [None], [None]
    for l, lev in enumerate(levels):
        b, p = lev[0], lev[1]
        dg = len(lev) > 2 and lev[2] == "dg"
        nq = None if n_q_points_1d is None else n_q_points_1d(p)
        if dg:
            op = LaplaceOperatorDG(b, p, nq, precision, penalty_kind=dg_penalty_kind,
                                   metric=dg_metric)
        else:
            op = LaplaceOperatorBK3(b, p, nq, coarse_precision if l == 0 else precision)
            if coefficient is not None:
                op.set_coefficient(coefficient(b, op.nq))
            if affine:
                op.use_affine_geometry()
        op.compute_diagonal()
        ops.append(op)
        if l == 0:
            assert not dg, "the coarsest level must be continuous"
            continue
        bc, pc = levels[l - 1][0], levels[l - 1][1]
        assert not (len(levels[l - 1]) > 2 and levels[l - 1][2] == "dg"), \
            "a DG level can only be the finest"
        same = (bc.nx, bc.ny, bc.nz) == (b.nx, b.ny, b.nz)
        if dg:
            assert same and pc == p, "a DG level sits on the CG level of the same box and degree"
            trs.append(ContinuousTransfer(b, p, precision))
        elif same:
            trs.append(PolynomialTransfer(b, pc, p, precision, device_weights=True))
        else:
            assert (2 * bc.nx, 2 * bc.ny, 2 * bc.nz) == (b.nx, b.ny, b.nz) and pc == p, \
                "consecutive levels must differ by one refinement or by the degree"
            trs.append(GeometricTransfer(bc, p, precision, prolongation_scale,
                                         device_weights=True))
        sms.append(PreconditionChebyshev(op, smoother_degree, smoothing_range,
                                         eig_cg_n_iterations))
    mg = VCycleMultigrid(ops, trs, sms, coarse, coarse_tol,
                         precision if outer_precision is None else outer_precision)
    return mg, ops, trs, sms


def ph_levels(fine_box, degree, n_h_levels, dg=False, p_sequence=None):
    """Level list of the reference's ph/DG drivers: Q1 on fine_box coarsened
    n_h_levels-1 times up to fine_box, then degrees 1 -> ... -> p on fine_box
    with p_levels = p, max(p-2,1), ... (tests/poisson_cube_bk3_fp64_dg/program.cc:294-300)
    or the given descending p_sequence (e.g. [6, 3, 1]), then (dg) DG(p) on
    fine_box."""
    boxes = coarsen_boxes(fine_box, n_h_levels)
    levels = [(b, 1) for b in boxes]
    ps = [degree]
    if p_sequence is not None:
        assert p_sequence[0] == degree and p_sequence[-1] == 1, "p_sequence runs from p down to 1"
        ps = list(p_sequence)
    while ps[-1] > 1:
        ps.append(max(ps[-1] - 2, 1))
    for p in reversed(ps[:-1]):
        levels.append((fine_box, p))
    if dg:
        levels.append((fine_box, degree, "dg"))
    return levels


def random_fourier_field(seed, n_modes=16, kmax=4):
    """Synthetic smooth field g(x) = sum_m a_m cos(2 pi k_m . x + phi_m): integer
    wave vectors k_m with max-norm 1..kmax, a_m ~ N(0, 0.5/|k_m|^2), phases
    uniform (numpy PCG64 with the seed).  Returns a callable X[n,3] -> g[n]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    modes = []
    for _ in range(n_modes):
        while True:
            k = rng.integers(-kmax, kmax + 1, 3)
            if np.abs(k).max() >= 1:
                break
        modes.append((k, rng.normal(0.0, np.sqrt(0.5 / float(k @ k))), rng.uniform(0, 2 * np.pi)))

    def g(X):
        X = np.asarray(X, dtype=np.float64)
        out = np.zeros(X.shape[0])
        for k, a, phi in modes:
            out += a * np.cos(2 * np.pi * (X @ k) + phi)
        return out

    return g


def quadrature_points(b, n_q_points_1d):
    """Physical Gauss points of the box's trilinear cells, [n_cells * nq^3, 3]
    in caller cell order, q-points lexicographic (x fastest) -- the layout of
    kappa for LaplaceOperatorBK3.set_coefficient (host-side set-up helper)."""
    nq = n_q_points_1d
    x, _ = np.polynomial.legendre.leggauss(nq)
    x = 0.5 * (x + 1.0)
    cv = box_cell_vertices(b).reshape(-1, 8, 3)
    xi = np.stack(np.meshgrid(x, x, x, indexing="ij"), -1)[..., ::-1].reshape(-1, 3)  # [q][xyz]
    N = np.empty((xi.shape[0], 8))
    for v in range(8):
        f = [xi[:, d] if (v >> d) & 1 else 1.0 - xi[:, d] for d in range(3)]
        N[:, v] = f[0] * f[1] * f[2]
    return np.einsum("qv,cvd->cqd", N, cv).reshape(-1, 3)


class CoefficientPoisson:
    """-div(kappa grad u) = f with kappa = exp(g) on a box: MG-CG forward and
    adjoint solves and the gradient dJ/dg at the quadrature points.

    The outer FP64 operator carries kappa = exp(g_q) at the fine Gauss points
    (any g_q); the V-cycle preconditioner (h-levels 2^k..fine, Chebyshev,
    coarse PCG, levels in `level_precision`) carries exp(g_smooth(x))
    evaluated at every level's Gauss points and stays fixed, so perturbing
    g_q needs no re-setup.  The gradient of J(u) for dJ/du = w is
    -kappa_q (grad u)^T G_q (grad lambda) with A lambda = w (A symmetric)."""

    def __init__(self, fine_box, degree, n_levels, g_smooth, level_precision=FP32,
                 n_q_points_1d=None, f_kind=0, rel_tol=1e-10, max_it=500):
        import torch
        self.box, self.degree = fine_box, degree
        self.nq = degree + 1 if n_q_points_1d is None else n_q_points_1d
        self.rel_tol, self.max_it = rel_tol, max_it

        def kappa_of(b, nq):
            return torch.tensor(np.exp(g_smooth(quadrature_points(b, nq))), device="cuda")

        self.mg, self.levels, _, _, _ = geometric_hierarchy(
            fine_box, degree, n_levels, self.nq, level_precision, outer_precision=FP64,
            coefficient=kappa_of)
        self.op = LaplaceOperatorBK3(fine_box, degree, self.nq, FP64)
        self.X = quadrature_points(fine_box, self.nq)
        self.b = self.op.assemble_rhs(f_kind)
        self.iterations = []

    def set_g(self, g_q):
        """g_q: CUDA float64 tensor [n_cells * nq^3]; kappa = exp(g_q)."""
        import torch
        self.op.set_coefficient(torch.exp(g_q).contiguous())

    def solve(self, rhs=None):
        x = self.op.initialize_dof_vector()
        its, _, _ = solve_cg(self.op, self.b if rhs is None else rhs, x, PRECOND_MG, self.mg,
                             self.rel_tol, self.max_it)
        self.iterations.append(its)
        return x

    def gradient(self, u, lam):
        """dJ/dg_q for the current kappa (forward u, adjoint lambda)."""
        import torch
        out = torch.zeros(self.op.n_cells * self.nq ** 3, dtype=torch.float64, device="cuda")
        self.op.gradient_coefficient(u, lam, out)
        return out


def coarsen_boxes(fine_box, n_levels):
    boxes = [fine_box]
    for _ in range(n_levels - 1):
        b = boxes[0]
        assert b.nx % 2 == 0 and b.ny % 2 == 0 and b.nz % 2 == 0, "box not coarsenable"
        boxes.insert(0, box(b.nx // 2, b.ny // 2, b.nz // 2, tuple(b.lo), tuple(b.hi), b.deform,
                            b.amplitude))
    return boxes


def geometric_hierarchy(fine_box, degree, n_levels, n_q_points_1d=None, precision=FP64,
                        smoother_degree=3, smoothing_range=15.0, eig_cg_n_iterations=10,
                        coarse=COARSE_PCG, coarse_tol=1e-12, prolongation_scale=1.0,
                        outer_precision=None, coefficient=None, coarse_precision=None,
                        affine=False):
    """h-only hierarchy at fixed degree (source/geometric_multigrid_bk3/program.cc):
    level 0 has fine_box coarsened n_levels-1 times."""
    boxes = coarsen_boxes(fine_box, n_levels)
    nqf = None if n_q_points_1d is None else (lambda p: n_q_points_1d)
    mg, ops, trs, sms = build_hierarchy([(b, degree) for b in boxes], nqf, precision,
                                        smoother_degree, smoothing_range, eig_cg_n_iterations,
                                        coarse, coarse_tol, prolongation_scale, outer_precision,
                                        coefficient=coefficient, coarse_precision=coarse_precision,
                                        affine=affine)
    return mg, ops, trs, sms, boxes
