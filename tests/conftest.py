"""Shared test setup: import paths, the `gpu` marker, golden-fixture loading."""
import glob
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "graspan-g_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP closure engine)")


def golden_files():
    return sorted(glob.glob(os.path.join(GOLDEN, "*.json")))


def load_golden(path):
    with open(path) as f:
        d = json.load(f)
    e = np.asarray(d["edges"], dtype=np.int64).reshape(-1, 3)
    d["src"] = e[:, 0].astype(np.uint32)
    d["dst"] = e[:, 1].astype(np.uint32)
    d["lab"] = e[:, 2].astype(np.uint8)
    d["rules_arr"] = np.asarray(d["rules"], dtype=np.int32).reshape(-1, 3)
    d["expected_arr"] = np.asarray(d["expected"], dtype=np.uint32).reshape(-1, 3)
    d["name"] = os.path.basename(path)
    return d


@pytest.fixture(scope="session")
def built():
    """Builds the oracle (CPU) once per session."""
    import oracle
    oracle.build()
    return True
