"""Shared fixtures: the package (directory wavi.jl_amd, imported as
wavi_jl_amd), the CPU oracle, and the `gpu` marker."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def load_package():
    if "wavi_jl_amd" in sys.modules:
        return sys.modules["wavi_jl_amd"]
    pkg_dir = os.path.join(ROOT, "wavi.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "wavi_jl_amd", os.path.join(pkg_dir, "__init__.py"), submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["wavi_jl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def W():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.load()
    return pyoracle
