import importlib.util
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
PKG_DIR = os.path.join(REPO, "distributed-lb-ib_amd")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


if PKG_DIR not in sys.path:
    # the package directory's name is not a Python identifier; its modules
    # (lbmib, lbmib_configs) are imported from the directory itself
    sys.path.insert(0, PKG_DIR)


def load_pkg_module(name):
    return importlib.import_module(name)


@pytest.fixture(scope="session")
def lbmib():
    return load_pkg_module("lbmib")


@pytest.fixture(scope="session")
def configs():
    return load_pkg_module("lbmib_configs")


@pytest.fixture(scope="session")
def fixtures_meta():
    with open(os.path.join(GOLDEN, "fixtures.json")) as fh:
        return json.load(fh)


def read_fiber_dump(path):
    """Reads the reference's binary fiber dump (lbm_ib.c:604-646):
    {int32 i, int32 j, 12 x float64} records.  Returns (ij [m,2], vals [m,12])."""
    raw = np.fromfile(path, dtype=np.uint8).reshape(-1, 104)
    ij = raw[:, :8].copy().view(np.int32).reshape(-1, 2)
    vals = raw[:, 8:].copy().view(np.float64).reshape(-1, 12)
    return ij, vals


def reference_sheet_xyz(meta):
    """Positions gen_fiber_shape produces for a fixture run (lbm_ib.c:899-906)."""
    rows, cols = meta["rows"], meta["cols"]
    w, h = meta["width"], meta["height"]
    x0, y0, z0 = meta["origin"]
    zd, yd = w / (cols - 1), h / (rows - 1)
    xyz = np.zeros((rows, cols, 3))
    for r in range(rows):
        for c in range(cols):
            xyz[r, c] = (x0, yd * r + y0, zd * c + z0)
    return xyz
