import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as ge  # noqa: E402

ge.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pytest_collection_finish(session):
    """torch ships a HIP runtime of its own beside the one the library links;
    in a process where the library reaches the GPU first, torch's first GPU
    call reports no device. GPU tests build eigenvector blocks with torch, so
    torch's runtime is started before any test runs."""
    if any(item.get_closest_marker("gpu") for item in session.items):
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()


@pytest.fixture(scope="session")
def engine():
    from gpu_intensities_amd import binding
    eng = binding.Engine(0)
    yield eng
    eng.close()
