import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: large inputs (full benchmark sizes)")


def has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    import pbbsbench_amd as pb
    c = pb.Context(0)
    yield c
    c.close()


def golden_cases():
    import json
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        man = json.load(f)
    return sorted(man.items())
