import glob
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "par-cd_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def golden(pattern):
    """Sorted golden fixture paths matching pattern (under tests/golden)."""
    return sorted(glob.glob(os.path.join(GOLDEN, pattern)))


def load(path):
    return dict(np.load(path, allow_pickle=False))


@pytest.fixture(scope="session")
def golden_graphs():
    import oracle_lib as O
    out = {}
    for p in golden("graph_*.npz"):
        name = os.path.basename(p)[6:-4]
        d = load(p)
        out[name] = O.parse_pcd(d["pcd"].tobytes()) + (d,)
    return out
