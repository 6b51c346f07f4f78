import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ctx():
    """One r2 context for the GPU tests (fails loudly without a HIP device)."""
    import renderer2_amd as r2

    c = r2.Context(0)
    yield c
    c.close()
