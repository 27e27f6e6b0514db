import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle.oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def fm():
    import fast_mul_ntt_amd

    fast_mul_ntt_amd.lib()
    return fast_mul_ntt_amd
