import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

FIXTURES = ROOT / "tests" / "fixtures"
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def experiments():
    import json
    return json.loads((FIXTURES / "experiments.json").read_text())


def load_fixture(stiffness):
    import numpy as np
    return dict(np.load(FIXTURES / f"fixture_{stiffness:.0e}.npz"))
