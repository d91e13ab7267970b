import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def built():
    """Build the oracle (and the HIP library when missing) once per session. On the GPU box
    the prebuilt in-tree libraries are used as shipped."""
    from oracle import build as ob

    if not os.path.exists(os.path.join(REPO, "oracle", "lib", "libmg_oracle.so")):
        ob.build_oracle()
    return True
