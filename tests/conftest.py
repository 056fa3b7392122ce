import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "sci-solver_level-set_amd")


def _load_package():
    if "levelset_amd" in sys.modules:
        return sys.modules["levelset_amd"]
    spec = importlib.util.spec_from_file_location(
        "levelset_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["levelset_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_package()
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def mlp_oracle():
    import mlp_oracle
    return mlp_oracle
