import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkm_hydro.so)")
    config.addinivalue_line("markers", "slow: long CPU oracle run")


@pytest.fixture(scope="session")
def oracle():
    from oracle import km_oracle
    km_oracle.lib()
    return km_oracle


@pytest.fixture(scope="session")
def km():
    import kamayan_amd
    kamayan_amd.lib()
    return kamayan_amd


@pytest.fixture(scope="session")
def gpu(km):
    if km.device_count() < 1:
        pytest.fail("GPU test collected but no GPU is visible")
    return km
