import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


# The batch autotuner keeps its choices in a file; the tests use a file of
# their own so that every session times afresh and nothing outside the run
# is written.
if "PGLD_TUNE_CACHE" not in os.environ:
    import tempfile
    os.environ["PGLD_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="pgld_tune_"), "batch_tune.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libpgld.so")
    config.addinivalue_line("markers", "slow: long-running")
