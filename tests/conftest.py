"""Shared pytest configuration.

Markers:
  gpu -- needs an MI355X (HIP device); parity tests through the C ABI.
Tests without the marker run on the CPU: the oracle against the reference's
golden vectors, host logic, and that the HIP library loads and exports every
symbol its header declares.
"""
import glob
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


def _ensure_oracle():
    if not glob.glob(os.path.join(ROOT, "oracle", "_oracle*.so")):
        from oracle import build_oracle
        build_oracle.build()
    from oracle import _oracle
    return _oracle


@pytest.fixture(scope="session")
def oracle_mod():
    return _ensure_oracle()


def oracle_sim(mod, dim, prec, origin, end, cell_size, shape_function, dt, ppc):
    cls = getattr(mod, f"Sim{dim}D_{prec}")
    return cls(origin, end, cell_size, shape_function, dt, ppc)
