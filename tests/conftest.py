import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

from oracle_lib import ORACLE_PATH, REF_PATH, Oracle, RefBuild  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP library")


@pytest.fixture(scope="session")
def orc():
    if not os.path.exists(ORACLE_PATH):
        pytest.fail(f"oracle not built: {ORACLE_PATH} (run __graft_entry__.build())")
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_PATH):
        pytest.skip("reference build oracle/_ref not present (needs the reference checkout at build time)")
    return RefBuild()


@pytest.fixture(scope="session")
def ex():
    """One library context on GPU 0 for the whole GPU test session."""
    import ozablas_amd as oz

    e = oz.HipExecutor(0)
    yield e
    e.close()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
