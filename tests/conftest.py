import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# every batch load in the tests times the SpMM shapes itself, unless a test
# names a tuning file of its own (the batch autotuner's persisted winners)
os.environ.setdefault("HECMW33_TUNE_CACHE", "off")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def has_gpu():
    import frontistr_amd as fa
    return fa.lib().hecmw33_device_count() > 0
