import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "goldens.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: takes more than ~20 s on the CPU")


@pytest.fixture(scope="session")
def goldens():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx():
    import signed_triangle_counting_amd as stc
    c = stc.Counter(0)
    # "learned" in every GPU test means the shipped model, not the fixed rule
    # it falls back to
    assert c.model_status == 0, c.last_error()
    yield c
    c.close()
