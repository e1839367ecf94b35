import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdisco_hip.so)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import bind
    if not os.path.exists(bind.LIB_PATH):
        bind.build()
    return bind.lib()
