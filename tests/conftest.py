"""Shared fixtures.  `gpu` marks tests that need an MI355X (run with -m gpu);
everything else runs on the CPU (oracle pins, host logic, ABI checks)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "3dcompactsolver_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X); run with -m gpu")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def hiplib():
    """The product library.  It must have been built (build() in
    __graft_entry__); a missing library is an error, not a skip."""
    import cs3d
    return cs3d.lib()
