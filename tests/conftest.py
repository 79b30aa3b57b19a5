"""Test configuration: the ``gpu`` marker and import paths.

``-m "not gpu"`` runs on any machine (oracle pinning, host logic, the C ABI's
exports and argument checks).  ``-m gpu`` needs an MI355X and compares the HIP
engine with the oracle through the C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); compares the engine with the oracle")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle
