import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "elvibrot-tnumtana_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libevrhip.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def evr():
    import evrhip
    evrhip._lib.load()
    return evrhip


def hh_potential(q, lam=0.111803):
    """Henon-Heiles potential of the reference's 6D/21D working tests
    (Working_tests/MPI_tests/6D_Davidson_S0/sub_system_HenonHeiles.f:33-48)."""
    D = q.shape[1]
    V = 0.5 * (q ** 2).sum(1)
    for i in range(D - 1):
        V = V + lam * (q[:, i] ** 2 * q[:, i + 1] - q[:, i + 1] ** 3 / 3.0)
    return V
