import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def load_csv(name, ncols=None):
    """Rows of a fixture CSV split on ',' / ' ' like read_file (FileReader.cc:54-83)."""
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            toks = [t for t in line.replace(",", " ").split() if t]
            rows.append([int(t) for t in toks])
    return rows
