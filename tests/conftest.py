import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def g1():
    """The reference's g1 fixture (13 weighted directed edges, ids 1..8) and its known answers."""
    with open(os.path.join(ROOT, "tests", "golden", "g1_kat.json")) as f:
        k = json.load(f)
    k["records_w"] = np.array(k["edges_weighted"], dtype=np.uint32)
    k["records"] = np.ascontiguousarray(k["records_w"][:, :2])
    return k


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    import la3_amd
    ctx = la3_amd.Context(0)
    yield ctx
    ctx.close()
