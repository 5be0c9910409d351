import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "flowvc-1_amd"))
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle_built():
    """Build the oracle (checker) once per session; it is plain gcc."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], capture_output=True, text=True)
    if r.returncode:
        pytest.fail("oracle build failed:\n" + r.stdout + r.stderr)
    return os.path.join(REPO, "oracle")


@pytest.fixture(scope="session")
def host_built():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "flowvc-1_amd"), "libfvc_host.so"],
                       capture_output=True, text=True)
    if r.returncode:
        pytest.fail("host library build failed:\n" + r.stdout + r.stderr)
    return True
