import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, HERE)
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def golden():
    return load_golden


def meta(d, key):
    v = d["meta_" + key]
    return v.item()


def product():
    """The product's ctypes binding (md-bench_amd/mdbench_gfx950.py)."""
    import importlib.util
    path = os.path.join(REPO, "md-bench_amd", "mdbench_gfx950.py")
    spec = importlib.util.spec_from_file_location("mdbench_gfx950", path)
    if "mdbench_gfx950" in sys.modules:
        return sys.modules["mdbench_gfx950"]
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mdbench_gfx950"] = mod
    spec.loader.exec_module(mod)
    return mod


def surrogate_module():
    """The surrogate's host module (md-bench_amd/eam_surrogate.py)."""
    import importlib.util
    product()
    if "eam_surrogate" in sys.modules:
        return sys.modules["eam_surrogate"]
    path = os.path.join(REPO, "md-bench_amd", "eam_surrogate.py")
    spec = importlib.util.spec_from_file_location("eam_surrogate", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["eam_surrogate"] = mod
    spec.loader.exec_module(mod)
    return mod
