"""Shared test setup: the `gpu` marker, paths, and loaders."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "flame-fractal-renderer_amd")
FLAMES = os.path.join(PKG, "flames")
GOLDEN = os.path.join(ROOT, "tests", "golden")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    # build only when nothing is built (a machine without the toolchain runs
    # what was built before)
    need = [os.path.join(PKG, "libffr_hip.so"), os.path.join(PKG, "ffr-buf"),
            os.path.join(ROOT, "oracle", "liboracle.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build(force=False)


def read_flame(name: str) -> str:
    with open(os.path.join(FLAMES, name)) as f:
        return f.read()


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ffr_mod():
    import ffr
    ffr.lib()
    return ffr
