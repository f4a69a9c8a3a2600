import os
import sys

# One BLAS thread for the oracle, as the reference runs (setupenv.sh:6-8):
# its results then do not depend on the machine's core count, and MKL's
# threaded LAPACK never shares the process with torch's OpenMP runtime
# (with both active, the oracle's dsyevr returned garbage).
os.environ["MKL_NUM_THREADS"] = "1"

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test (oracle end-to-end)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py
    oracle_py.lib()
    return oracle_py
