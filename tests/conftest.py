import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_built():
    """Build the oracle and the HIP library when either is missing (hipcc cross-compiles for
    gfx950 without a GPU). On the GPU box the prebuilt in-tree libraries are used as shipped."""
    from morphogenesis_amd import build as mb
    from oracle import build as ob

    if not os.path.exists(os.path.join(REPO, "oracle", "lib", "libmg_oracle.so")):
        ob.build_oracle()
    if not os.path.exists(os.path.join(REPO, "morphogenesis_amd", "lib", "libmg_sph.so")):
        mb.build_library()
    if not os.path.exists(os.path.join(REPO, "tests", "hip", "lib", "librounding_check.so")):
        mb.build_rounding_check()


def pytest_sessionstart(session):
    # scenes derive their parameters and grids through the library's host entry points
    _ensure_built()


@pytest.fixture(scope="session")
def built():
    _ensure_built()
    return True
