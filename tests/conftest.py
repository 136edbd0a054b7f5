import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "ref: needs the reference routines built in oracle/_ref")


def pytest_collection_modifyitems(config, items):
    from oracle import pyoracle
    if not pyoracle.ref_available():
        skip = pytest.mark.skip(reason="oracle/_ref not built (reference tree absent)")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
