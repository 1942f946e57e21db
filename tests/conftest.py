"""Shared pytest setup: the `gpu` marker and import paths.

Tests marked ``gpu`` need an MI355X (gfx950) and the built HIP library; all
other tests run on the CPU (oracle against the reference's golden vectors,
host logic, and that the C-ABI library loads and exports its symbols)."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cuda-spatial-filtering_amd")
for p in (REPO, PKG, os.path.join(REPO, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
