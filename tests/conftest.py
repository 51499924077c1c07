"""Test configuration: markers and import paths.

`gpu` tests need a real gfx950 device and the built HIP library; everything
else runs on the CPU.  The oracle lives in oracle/ (test infrastructure),
the product modules in first-knowns_amd/.
"""
import json
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for sub in ("oracle", "first-knowns_amd"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers",
                            "slow: long CPU-oracle reproductions (FK_SLOW=1)")


@pytest.fixture(scope="session")
def golden():
    path = ROOT / "tests" / "golden" / "reference_vectors.json"
    return json.loads(path.read_text())


def big(v):
    """Fixture integers >= 2^53 are stored as decimal strings."""
    return int(v)
