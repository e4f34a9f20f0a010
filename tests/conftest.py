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
