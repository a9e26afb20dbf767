"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import ctypes
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "dphpc-project_amd")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


def load_pkg():
    name = "dphpc_project_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    """The oracle's C restatement (test infrastructure only)."""
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
    lib = ctypes.CDLL(ORACLE_LIB)
    lib.orc_mcpi_hits.restype = ctypes.c_int64
    lib.orc_npb_is_run.restype = ctypes.c_int64
    lib.orc_npb_is_iteration.restype = ctypes.c_int64
    lib.orc_heap_check.restype = ctypes.c_int64
    return lib


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch
