import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; calls the engine through the C-ABI")


@pytest.fixture(scope="session")
def ctx():
    import cuminlpsearch_amd as cmi

    return cmi.Context(0)
