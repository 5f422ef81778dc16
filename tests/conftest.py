"""Test configuration: registers the `gpu` marker and loads the package.

The package directory is named `symmetriceigenproblem.jl_amd`, which is not a valid
Python identifier, so it is loaded here under the module name `sep_amd`.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "symmetriceigenproblem.jl_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    if "sep_amd" in sys.modules:
        return sys.modules["sep_amd"]
    spec = importlib.util.spec_from_file_location(
        "sep_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sep_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def sep():
    return load_package()


@pytest.fixture(scope="session")
def handle(sep):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    h = sep.Handle(0)
    yield h
    h.close()
