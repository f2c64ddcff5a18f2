"""Shared pytest setup: the `gpu` marker, import paths and library fixtures.

CPU tests (-m "not gpu") cover the oracle against closed forms, brute force
and invariants, the host-side logic of the product library, and that the
C-ABI library loads and exports every declared symbol.  GPU tests call the
HIP path through the C ABI and compare it with the oracle.
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PY = os.path.join(REPO, "sph-fluid-simulator_amd", "python")
for p in (PKG_PY, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def orc32():
    from oracle_binding import Oracle
    return Oracle(32)


@pytest.fixture(scope="session")
def orc64():
    from oracle_binding import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def sph():
    import sph_amd
    sph_amd.load()
    return sph_amd
