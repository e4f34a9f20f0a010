"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle restates the reference algorithms of gshubhamk/portable-multigrid in
plain C (see pmg_oracle.h for the per-routine citations and what pins each
one).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
may import this package, and only as the checker or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libpmg_oracle.so")
INVALID = 0xFFFFFFFF

_lib = None


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


class Mesh(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("deform", ctypes.c_int), ("amplitude", ctypes.c_double)]


def mesh(nx, ny=None, nz=None, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), deform=0, amplitude=0.0):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    m = Mesh()
    m.nx, m.ny, m.nz = nx, ny, nz
    m.lo[:] = lo
    m.hi[:] = hi
    m.deform = deform
    m.amplitude = amplitude
    return m


class MGParams(ctypes.Structure):
    _fields_ = [("n_levels", ctypes.c_int), ("n_threads", ctypes.c_int),
                ("cheby_degree", ctypes.c_int), ("smoothing_range", ctypes.c_double),
                ("eig_cg_n_iterations", ctypes.c_int), ("coarse_kind", ctypes.c_int),
                ("coarse_tol", ctypes.c_double), ("h_prolong_scale", ctypes.c_double),
                ("dg_top", ctypes.c_int), ("dg_penalty_kind", ctypes.c_int),
                ("dg_metric", ctypes.c_int)]


_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_MP = ctypes.POINTER(Mesh)
_L = ctypes.c_long
_i = ctypes.c_int
_d = ctypes.c_double
_vp = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "pmo_gauss_legendre": (None, [_i, _D, _D]),
            "pmo_gauss_lobatto": (None, [_i, _D]),
            "pmo_lagrange": (None, [_i, _D, _d, _vp, _vp]),
            "pmo_shape_matrices": (None, [_i, _i, _D, _D]),
            "pmo_shape_gradients": (None, [_i, _i, _D]),
            "pmo_n_cells": (_L, [_MP]),
            "pmo_n_dofs_cg": (_L, [_MP, _i]),
            "pmo_vertices": (None, [_MP, _D]),
            "pmo_cell_vertices": (None, [_MP, _D]),
            "pmo_dofmap_cg": (None, [_MP, _i, _i, _U]),
            "pmo_boundary_dofs_cg": (_L, [_MP, _i, _vp]),
            "pmo_geometry": (None, [_MP, _i, _vp, _vp, _vp]),
            "pmo_bk3_cells": (None, [_i, _i, _L, _D, _D, _D, _U, _D, _D, _i, _vp]),
            "pmo_bk3_vmult": (None, [_i, _i, _L, _D, _D, _D, _U, _L, _U, _L, _D, _D, _i, _vp]),
            "pmo_bk3_vmult_subdomain": (None, [_i, _i, _L, _D, _D, _D, _U, _L, _U, _L, _vp, _L,
                                                _D, _D]),
            "pmo_bk3_element_matrices": (None, [_i, _i, _L, _D, _D]),
            "pmo_bk3_diagonal": (None, [_i, _i, _L, _D, _U, _L, _U, _L, _D]),
            "pmo_cell_colors": (None, [_MP, _I]),
            "pmo_rhs": (None, [_MP, _i, _i, _i, _U, _L, _D]),
            "pmo_l2_error": (_d, [_MP, _i, _i, _i, _U, _D]),
            "pmo_scale_G": (None, [_L, _i, _D, _D, _D]),
            "pmo_grad_kappa": (None, [_i, _i, _L, _D, _D, _D, _D, _U, _D, _D, _D]),
            "pmo_h_matrix": (None, [_i, _D]),
            "pmo_p_matrix": (None, [_i, _i, _D]),
            "pmo_h_maps": (None, [_MP, _i, _U, _U, _D]),
            "pmo_p_maps": (None, [_MP, _i, _i, _U, _U, _D]),
            "pmo_prolongate": (None, [_i, _i, _L, _D, _U, _U, _D, _d, _D, _D]),
            "pmo_restrict": (None, [_i, _i, _L, _D, _U, _U, _D, _D, _D]),
            "pmo_mg_create": (_vp, [_i, _MP, _I, _vp, ctypes.POINTER(MGParams)]),
            "pmo_mg_destroy": (None, [_vp]),
            "pmo_mg_n_dofs": (_L, [_vp, _i]),
            "pmo_mg_vmult": (None, [_vp, _D, _D]),
            "pmo_mg_level_vmult": (None, [_vp, _i, _D, _D]),
            "pmo_mg_smoother_info": (None, [_vp, _i, ctypes.POINTER(_d), ctypes.POINTER(_d),
                                            ctypes.POINTER(_i)]),
            "pmo_mg_smooth": (None, [_vp, _i, _D, _D, _i]),
            "pmo_pcg": (_i, [_vp, _i, _i, _D, _D, _d, _i, ctypes.POINTER(_d),
                             ctypes.POINTER(_d)]),
            "pmo_tridiag_eigs": (None, [_i, _D, _D, _D]),
            "pmo_dg_n_dofs": (_L, [_MP, _i]),
            "pmo_dg_penalty": (None, [_MP, _i, _i, _D]),
            "pmo_dg_vmult": (None, [_MP, _i, _i, _D, _i, _i, _D, _D]),
            "pmo_dg_cg_prolongate": (None, [_i, _L, _U, _D, _D]),
            "pmo_dg_cg_restrict": (None, [_i, _L, _U, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ 1D data
def gauss_legendre(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().pmo_gauss_legendre(n, x, w)
    return x, w


def gauss_lobatto(n):
    x = np.zeros(n)
    lib().pmo_gauss_lobatto(n, x)
    return x


def lagrange(nodes, x):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    v, d = np.zeros(len(nodes)), np.zeros(len(nodes))
    lib().pmo_lagrange(len(nodes), nodes, float(x), _ptr(v), _ptr(d))
    return v, d


def shape_matrices(p, nq):
    S, D = np.zeros((p + 1) * nq), np.zeros(nq * nq)
    lib().pmo_shape_matrices(p, nq, S, D)
    return S, D


def shape_gradients(p, nq):
    Sd = np.zeros((p + 1) * nq)
    lib().pmo_shape_gradients(p, nq, Sd)
    return Sd


# ------------------------------------------------------------------ mesh
def n_cells(m):
    return lib().pmo_n_cells(ctypes.byref(m))


def n_dofs(m, p):
    return lib().pmo_n_dofs_cg(ctypes.byref(m), p)


def cell_vertices(m):
    cv = np.zeros(n_cells(m) * 24)
    lib().pmo_cell_vertices(ctypes.byref(m), cv)
    return cv


def dofmap(m, p, dirichlet=True):
    out = np.zeros(n_cells(m) * (p + 1) ** 3, dtype=np.uint32)
    lib().pmo_dofmap_cg(ctypes.byref(m), p, int(dirichlet), out)
    return out


def boundary_dofs(m, p):
    n = lib().pmo_boundary_dofs_cg(ctypes.byref(m), p, None)
    out = np.zeros(n, dtype=np.uint32)
    lib().pmo_boundary_dofs_cg(ctypes.byref(m), p, _ptr(out))
    return out


def geometry(m, nq, want_jxw=False, want_x=False):
    C = n_cells(m)
    G = np.zeros(C * 6 * nq ** 3)
    J = np.zeros(C * nq ** 3) if want_jxw else None
    X = np.zeros(C * nq ** 3 * 3) if want_x else None
    lib().pmo_geometry(ctypes.byref(m), nq, _ptr(G), _ptr(J), _ptr(X))
    return G, J, X


def cell_colors(m):
    c = np.zeros(n_cells(m), dtype=np.int32)
    lib().pmo_cell_colors(ctypes.byref(m), c)
    return c


# ------------------------------------------------------------------ BK3
class LaplaceBK3:
    """Oracle counterpart of LaplaceOperatorBK3 on a structured mesh."""

    def __init__(self, m, p, nq=None, G=None, n_threads=1):
        self.m, self.p = m, p
        self.nq = p + 1 if nq is None else nq
        self.C = n_cells(m)
        self.N = n_dofs(m, p)
        self.S, self.D = shape_matrices(p, self.nq)
        self.G = geometry(m, self.nq)[0] if G is None else np.ascontiguousarray(G)
        self.map = dofmap(m, p, True)
        self.plain_map = dofmap(m, p, False)
        self.cons = boundary_dofs(m, p)
        self.colors = cell_colors(m)
        self.n_threads = n_threads

    def vmult(self, src, n_threads=None):
        dst = np.zeros(self.N)
        nt = self.n_threads if n_threads is None else n_threads
        lib().pmo_bk3_vmult(self.p, self.nq, self.C, self.S, self.D, self.G, self.map, self.N,
                            self.cons, len(self.cons), np.ascontiguousarray(src, np.float64),
                            dst, nt, _ptr(self.colors))
        return dst

    def vmult_subdomain(self, src, physical, cells=None):
        """vmult_neumann (cells None) or vmult_interface_cell_range (the listed
        cells) of the subdomain operator: plain map, physical DoFs eliminated."""
        dst = np.zeros(self.N)
        phys = np.ascontiguousarray(physical, dtype=np.uint32)
        sel = None if cells is None else np.ascontiguousarray(cells, dtype=np.uint32)
        lib().pmo_bk3_vmult_subdomain(self.p, self.nq, self.C, self.S, self.D, self.G,
                                      self.plain_map, self.N, phys, len(phys), _ptr(sel),
                                      0 if sel is None else len(sel),
                                      np.ascontiguousarray(src, np.float64), dst)
        return dst

    def cells_add(self, src, dst, n_threads=1):
        lib().pmo_bk3_cells(self.p, self.nq, self.C, self.S, self.D, self.G, self.map,
                            np.ascontiguousarray(src, np.float64), dst, n_threads,
                            _ptr(self.colors))
        return dst

    def diagonal(self):
        d = np.zeros(self.N)
        lib().pmo_bk3_diagonal(self.p, self.nq, self.C, self.G, self.map, self.N, self.cons,
                               len(self.cons), d)
        return d

    def element_matrices(self):
        nm3 = (self.p + 1) ** 3
        K = np.zeros(self.C * nm3 * nm3)
        lib().pmo_bk3_element_matrices(self.p, self.nq, self.C, self.G, K)
        return K.reshape(self.C, nm3, nm3)

    def rhs(self, f_kind, nq=None):
        b = np.zeros(self.N)
        lib().pmo_rhs(ctypes.byref(self.m), self.p, self.p + 1 if nq is None else nq, f_kind,
                      self.map, self.N, b)
        return b

    def l2_error(self, u, u_kind=1, nq=None):
        return lib().pmo_l2_error(ctypes.byref(self.m), self.p, self.p + 2 if nq is None else nq,
                                  u_kind, self.plain_map, np.ascontiguousarray(u, np.float64))

    def grad_kappa(self, kappa, u, lam):
        out = np.zeros(self.C * self.nq ** 3)
        lib().pmo_grad_kappa(self.p, self.nq, self.C, self.S, self.D, self.G,
                             np.ascontiguousarray(kappa, np.float64), self.map,
                             np.ascontiguousarray(u, np.float64),
                             np.ascontiguousarray(lam, np.float64), out)
        return out


class LaplaceDG:
    """Oracle counterpart of LaplaceOperatorDG (SIP) on a structured mesh.

    DoFs are cell-major; sigma is [C][6] (per cell face, the same value seen
    from both sides of an inner face); metric 0 uses J^{-1}n, 1 J^{-T}n."""

    def __init__(self, m, p, nq=None, penalty_kind=0, sigma=None, metric=0, neumann_mask=0):
        self.m, self.p = m, p
        self.nq = p + 1 if nq is None else nq
        self.C = n_cells(m)
        self.N = self.C * (p + 1) ** 3
        self.metric, self.neumann_mask = metric, neumann_mask
        if sigma is None:
            sigma = np.zeros(6 * self.C)
            lib().pmo_dg_penalty(ctypes.byref(m), p, penalty_kind, sigma)
        self.sigma = np.ascontiguousarray(sigma, np.float64)
        self.map = np.arange(self.N, dtype=np.uint32)

    def vmult(self, src):
        dst = np.zeros(self.N)
        lib().pmo_dg_vmult(ctypes.byref(self.m), self.p, self.nq, self.sigma, self.metric,
                           self.neumann_mask, np.ascontiguousarray(src, np.float64), dst)
        return dst

    def cell_diagonal(self):
        """Diagonal of the cell integrals only (what LaplaceOperatorDG uses)."""
        G = geometry(self.m, self.nq)[0]
        d = np.zeros(self.N)
        lib().pmo_bk3_diagonal(self.p, self.nq, self.C, G, self.map, self.N,
                               np.zeros(1, np.uint32), 0, d)
        return d

    def rhs(self, f_kind, nq=None):
        b = np.zeros(self.N)
        lib().pmo_rhs(ctypes.byref(self.m), self.p, self.p + 1 if nq is None else nq, f_kind,
                      self.map, self.N, b)
        return b

    def l2_error(self, u, u_kind=1, nq=None):
        return lib().pmo_l2_error(ctypes.byref(self.m), self.p, self.p + 2 if nq is None else nq,
                                  u_kind, self.map, np.ascontiguousarray(u, np.float64))


def dg_cg_prolongate(p, cg_map, src_cg, dst_dg):
    C = len(cg_map) // (p + 1) ** 3
    lib().pmo_dg_cg_prolongate(p, C, np.ascontiguousarray(cg_map, np.uint32),
                               np.ascontiguousarray(src_cg, np.float64), dst_dg)
    return dst_dg


def dg_cg_restrict(p, cg_map, src_dg, dst_cg):
    C = len(cg_map) // (p + 1) ** 3
    lib().pmo_dg_cg_restrict(p, C, np.ascontiguousarray(cg_map, np.uint32),
                             np.ascontiguousarray(src_dg, np.float64), dst_cg)
    return dst_cg


def scale_G(G, kappa, nq):
    C = len(kappa) // nq ** 3
    out = np.zeros_like(G)
    lib().pmo_scale_G(C, nq, np.ascontiguousarray(kappa, np.float64), G, out)
    return out


# ------------------------------------------------------------------ transfers
def h_matrix(p):
    M = np.zeros((p + 1) * (2 * p + 1))
    lib().pmo_h_matrix(p, M)
    return M


def p_matrix(pc, pf):
    M = np.zeros((pc + 1) * (pf + 1))
    lib().pmo_p_matrix(pc, pf, M)
    return M


class Transfer:
    """h (coarse mesh -> 2x refined, same p) or p (same mesh, pc -> pf) transfer."""

    def __init__(self, coarse_mesh, pc, pf=None, kind="h", h_prolong_scale=1.0):
        self.kind = kind
        C = n_cells(coarse_mesh)
        if kind == "h":
            self.nmc, self.nmf = pc + 1, 2 * pc + 1
            self.M = h_matrix(pc)
        else:
            self.nmc, self.nmf = pc + 1, pf + 1
            self.M = p_matrix(pc, pf)
        self.C = C
        self.map_c = np.zeros(C * self.nmc ** 3, dtype=np.uint32)
        self.map_f = np.zeros(C * self.nmf ** 3, dtype=np.uint32)
        self.w = np.zeros(C * self.nmf ** 3)
        if kind == "h":
            lib().pmo_h_maps(ctypes.byref(coarse_mesh), pc, self.map_c, self.map_f, self.w)
        else:
            lib().pmo_p_maps(ctypes.byref(coarse_mesh), pc, pf, self.map_c, self.map_f, self.w)
        self.scale = h_prolong_scale if kind == "h" else 1.0

    def prolongate_add(self, xc, yf):
        lib().pmo_prolongate(self.nmc, self.nmf, self.C, self.M, self.map_c, self.map_f, self.w,
                             self.scale, np.ascontiguousarray(xc, np.float64), yf)
        return yf

    def restrict_add(self, xf, yc):
        lib().pmo_restrict(self.nmc, self.nmf, self.C, self.M, self.map_c, self.map_f, self.w,
                           np.ascontiguousarray(xf, np.float64), yc)
        return yc


# ------------------------------------------------------------------ multigrid
class Multigrid:
    """Oracle V-cycle hierarchy: level l has mesh meshes[l] and degree degrees[l]."""

    def __init__(self, meshes, degrees, nqs=None, cheby_degree=3, smoothing_range=15.0,
                 eig_cg_n_iterations=10, coarse_kind=1, coarse_tol=1e-12, h_prolong_scale=1.0,
                 n_threads=1, dg_top=False, dg_penalty_kind=0, dg_metric=0):
        """dg_top: the finest level is SIP-DG on the mesh and degree of the level below."""
        L = len(meshes)
        self._meshes = (Mesh * L)(*meshes)
        self._ps = np.array(degrees, dtype=np.int32)
        self._nqs = None if nqs is None else np.array(nqs, dtype=np.int32)
        prm = MGParams(L, n_threads, cheby_degree, smoothing_range, eig_cg_n_iterations,
                       coarse_kind, coarse_tol, h_prolong_scale, int(dg_top), dg_penalty_kind,
                       dg_metric)
        self.h = lib().pmo_mg_create(L, self._meshes, self._ps, _ptr(self._nqs),
                                     ctypes.byref(prm))
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            lib().pmo_mg_destroy(self.h)
            self.h = None

    def n_dofs(self, level):
        return lib().pmo_mg_n_dofs(self.h, level)

    def vmult(self, src):
        dst = np.zeros(self.n_dofs(self.L - 1))
        lib().pmo_mg_vmult(self.h, dst, np.ascontiguousarray(src, np.float64))
        return dst

    def level_vmult(self, level, src):
        dst = np.zeros(self.n_dofs(level))
        lib().pmo_mg_level_vmult(self.h, level, dst, np.ascontiguousarray(src, np.float64))
        return dst

    def smoother_info(self, level):
        mx, mn, deg = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib().pmo_mg_smoother_info(self.h, level, ctypes.byref(mx), ctypes.byref(mn),
                                   ctypes.byref(deg))
        return mx.value, mn.value, deg.value

    def smooth(self, level, x, b, step):
        lib().pmo_mg_smooth(self.h, level, x, np.ascontiguousarray(b, np.float64), int(step))
        return x

    def pcg(self, b, level=None, precond=2, rel_tol=1e-10, max_it=1000, x0=None):
        level = self.L - 1 if level is None else level
        x = np.zeros(self.n_dofs(level)) if x0 is None else np.array(x0, dtype=np.float64)
        r0, r = ctypes.c_double(), ctypes.c_double()
        its = lib().pmo_pcg(self.h, level, precond, np.ascontiguousarray(b, np.float64), x,
                            rel_tol, max_it, ctypes.byref(r0), ctypes.byref(r))
        return x, its, r0.value, r.value


def tridiag_eigs(diag, off):
    n = len(diag)
    e = np.zeros(n)
    lib().pmo_tridiag_eigs(n, np.ascontiguousarray(diag, np.float64),
                           np.ascontiguousarray(np.append(off, 0.0), np.float64), e)
    return e
