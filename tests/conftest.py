"""Test configuration: the `gpu` marker, import paths and shared fixtures.

Tests marked `gpu` need an AMD GPU (they call libnq through its C ABI);
everything else runs on the CPU.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "n-queens_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def known():
    return json.loads((GOLDEN / "known.json").read_text())


@pytest.fixture(scope="session")
def digests():
    return json.loads((GOLDEN / "p5_digests.json").read_text())


@pytest.fixture(scope="session")
def p5_full():
    with np.load(GOLDEN / "p5_full.npz") as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def ctx():
    import nqueens
    c = nqueens.Context(0)
    yield c
    c.close()
