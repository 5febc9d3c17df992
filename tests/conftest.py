import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mind.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmind on cuda:0)")


def gpu_available() -> bool:
    from mind_amd import _abi

    return _abi.device_count() > 0


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    from mind_amd import _abi

    print(f"libmind: {_abi.LIB_PATH}; mechanisms: {_abi.build_info()}", flush=True)
    return True
