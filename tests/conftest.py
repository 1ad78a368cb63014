import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the HIP library on device 0 (GPU tests only)."""
    import qumirdb_amd
    qumirdb_amd.init_device(0)
    return qumirdb_amd


def pytest_runtest_logreport(report):
    """Prints a failure's report as soon as the test ends, so that it is
    kept even when a later test stops the run."""
    if report.failed and report.when == "call":
        sys.stdout.write("\n--- %s failed ---\n%s\n" % (report.nodeid, report.longreprtext[-4000:]))
        sys.stdout.flush()
