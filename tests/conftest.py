import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "cliquetrees.jl_amd")


def _load():
    if "cliquetrees_amd" in sys.modules:
        return sys.modules["cliquetrees_amd"]
    spec = importlib.util.spec_from_file_location(
        "cliquetrees_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["cliquetrees_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load()
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def ct():
    return _load()


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    return o
