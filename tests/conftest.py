import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


import pytest  # noqa: E402


@pytest.fixture(scope="session")
def ctx():
    """One gbk context for the whole GPU session (fails loudly without a GPU)."""
    from graphbrew_amd import gbk
    c = gbk.Context(0)
    yield c
    c.close()
