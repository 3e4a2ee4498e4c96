"""pytest configuration: the `gpu` marker and import paths.

Tests marked `gpu` need an MI355X; everything else runs on the CPU.
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PARENT = os.path.join(REPO, "differentialequations.jl_amd")
for p in (REPO, PKG_PARENT, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    return oracle_lib.load()
