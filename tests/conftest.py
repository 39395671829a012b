import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "2d-navier-stokes_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    # the ensemble's tuned-schedule cache of this test session, not the user's
    import tempfile
    os.environ.setdefault("NS2D_ENS_CACHE", os.path.join(tempfile.mkdtemp(prefix="ns2d_tests_"),
                                                         "ens_schedule.jsonl"))
