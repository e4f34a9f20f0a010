"""Shared pytest setup: the `gpu` marker, the package loader and small helpers."""
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_DIR = os.path.join(ROOT, "portable-multigrid_amd")


def load_pmg():
    """Import the package directory (its name is not a Python identifier)."""
    name = "portable_multigrid_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpmg.so on cuda:0)")


@pytest.fixture(scope="session")
def pmg():
    return load_pmg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def seeded_uniform(n, seed, lo=-1.0, hi=1.0):
    """U[lo,hi] from numpy's PCG64 with a fixed seed."""
    return np.random.Generator(np.random.PCG64(seed)).uniform(lo, hi, n)


def subdomain_sets(nx, ny, nz, p):
    """A box as one subdomain whose x = hi face is the interface to its
    neighbour (lexicographic Q_p numbering, x fastest).  Returns the physical
    boundary DoFs (on any other face, edges shared with the interface
    included), the interface-only DoFs, and the interface cells (a face on
    x = hi), as SubdomainDoFHandler classifies them."""
    n1 = np.array([p * nx + 1, p * ny + 1, p * nz + 1])
    d = np.arange(int(n1.prod()), dtype=np.int64)
    ix, iy, iz = d % n1[0], (d // n1[0]) % n1[1], d // (n1[0] * n1[1])
    physical = (ix == 0) | (iy == 0) | (iy == n1[1] - 1) | (iz == 0) | (iz == n1[2] - 1)
    interface = (ix == n1[0] - 1) & ~physical
    c = np.arange(nx * ny * nz, dtype=np.int64)
    cells = c[c % nx == nx - 1]
    return (d[physical].astype(np.uint32), d[interface].astype(np.uint32),
            cells.astype(np.uint32))
