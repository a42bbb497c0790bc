import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def reg_factory():
    """Registrar factory (GPU tests): each call returns a fresh engine with the given settings."""
    import zcmir_amd

    made = []

    def make(**settings):
        r = zcmir_amd.Registrar(**settings)
        made.append(r)
        return r

    yield make
    for r in made:
        r.close()
