import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcac_hip.so)")


@pytest.fixture(scope="session")
def orc():
    """The CPU FP64 oracle (built on first use)."""
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def cu_funcfl():
    from cac_code_amd.potential import johnson_cu_funcfl
    return johnson_cu_funcfl()


@pytest.fixture(scope="session")
def cu_tables(orc, cu_funcfl):
    return orc.eam_tables(cu_funcfl)
