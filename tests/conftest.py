import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "libosb_oracle.so")
HIP_SO = os.path.join(ROOT, "opensbli_amd", "libosb_hip.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU run, selected with --runslow")


def pytest_addoption(parser):
    parser.addoption("--runslow", action="store_true", default=False, help="run slow CPU tests")


def pytest_collection_modifyitems(config, items):
    if config.getoption("--runslow"):
        return
    skip = pytest.mark.skip(reason="slow; use --runslow")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def _ensure(path, target_dir):
    if not os.path.exists(path):
        subprocess.check_call(["make", "-C", target_dir], stdout=subprocess.DEVNULL)
    return path


@pytest.fixture(scope="session")
def oracle():
    from opensbli_amd import native as nat
    return nat.Library(_ensure(ORACLE_SO, os.path.join(ROOT, "oracle")))


@pytest.fixture(scope="session")
def hip():
    from opensbli_amd import native as nat
    _ensure(HIP_SO, os.path.join(ROOT, "opensbli_amd", "csrc"))
    return nat.load_hip()


@pytest.fixture(scope="session")
def hip_so():
    """Path of the built HIP library (built here if missing; loading needs no GPU)."""
    return _ensure(HIP_SO, os.path.join(ROOT, "opensbli_amd", "csrc"))
