import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def mt_golden():
    d = np.load(os.path.join(GOLDEN, "mt19937_64.npz"))
    seeds = [int(s) for s in d["seeds"]]
    return {"streams": dict(zip(seeds, d["streams"])), "kat": int(d["default_seed_draw_10000"][0])}


@pytest.fixture(scope="session")
def er():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import eigenrand_amd
    eigenrand_amd._lib.load()
    return eigenrand_amd
