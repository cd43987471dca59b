import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REFERENCE = "/root/reference"
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden():
    meta = json.load(open(os.path.join(GOLDEN, "reference_renders.json")))
    arrays = np.load(os.path.join(GOLDEN, "reference_renders.npz"))
    return meta, {k: arrays[k] for k in arrays.files}


def fixture_objects(meta_entry):
    """Objects of a reference fixture.  The fixtures were rendered by Lensed
    before its Sersic b(n) update (CHANGELOG.md:21,30), so their `sersic`
    is the `sersic-old` object."""
    return [(oid, "sersic-old" if name == "sersic" else name) for oid, name in meta_entry["objects"]]
