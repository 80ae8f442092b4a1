import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (select with -m gpu)")


def golden(name):
    """Load a committed fixture (npz: plain arrays only, no pickles)."""
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def meta():
    with open(os.path.join(GOLDEN, "ref_meta.json")) as fh:
        return json.load(fh)


def flush(msg):
    """Print and flush a progress line naming the case before a GPU step."""
    print(msg, flush=True)
