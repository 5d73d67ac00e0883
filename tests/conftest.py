import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "gpu-n-queens_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libnqm.so")


def golden_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def golden_npz(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


@pytest.fixture(scope="session")
def known_totals():
    return {int(k): v for k, v in golden_json("known_totals.json")["totals"].items()}


@pytest.fixture(scope="session")
def node_totals():
    return {int(k): v for k, v in golden_json("node_totals.json")["nodes"].items()}


@pytest.fixture(scope="session")
def ctx():
    import nqm
    c = nqm.Context(0)
    yield c
    c.close()
