import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_mesh():
    from lionbolt_amd import mesh
    return mesh.jittered_cube(3, a=2.0, seed=1001)


@pytest.fixture(scope="session")
def cube48k_mesh():
    from lionbolt_amd import mesh
    return mesh.jittered_cube(20, a=10.0, seed=1001)


@pytest.fixture(autouse=True, scope="module")
def _return_cached_device_memory():
    """After each test module, collect its operators and hand torch's cached device
    blocks back to HIP: the full-size tests cache over 150 GB in torch's
    allocator, which the library's own allocations in later modules then need."""
    yield
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        import gc
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
