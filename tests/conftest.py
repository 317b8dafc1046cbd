import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mtt-gpu-solver_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


@pytest.fixture(scope="session")
def mtt():
    # PyTorch ships its own HIP runtime; importing it before libmtt.so is
    # loaded makes the library bind to that same runtime in this process
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return _load("mtt", os.path.join(PKG, "mtt.py"))


@pytest.fixture(scope="session")
def solver_api(mtt):
    return _load("mtt_solver", os.path.join(PKG, "solver.py"))


@pytest.fixture(scope="session")
def oracle():
    return _load("oracle_ctypes", os.path.join(ORACLE, "oracle_ctypes.py"))


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    return torch.device("cuda:0")
