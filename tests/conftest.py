import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pfc-exp-gui_amd")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (parity checker)."""
    return _load("pfc_oracle_binding", os.path.join(ROOT, "oracle", "oracle.py"))


@pytest.fixture(scope="session")
def hip():
    """The product's ctypes binding; raises if libpfc_hip.so is not built."""
    mod = _load("pfc_hip", os.path.join(PKG, "pfc_hip.py"))
    mod.load()
    return mod
