import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, ORACLE_DIR)

import oracle as O  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (ROCm) to run")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(O.ORACLE_SO):
        subprocess.run(["make", "-C", ORACLE_DIR, "libfsoracle.so"], check=True, capture_output=True)
    return O.load_oracle()


@pytest.fixture(scope="session")
def ref():
    if not O.ref_available():
        pytest.skip("reference library oracle/_ref/libfsref.so not built here")
    return O.load_ref()


@pytest.fixture(scope="session", params=["oracle", "ref"])
def impl(request, oracle):
    """The FP64 restatement, and the compiled reference where it exists."""
    if request.param == "oracle":
        return oracle
    if not O.ref_available():
        pytest.skip("reference library not built here")
    return O.load_ref()
