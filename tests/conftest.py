import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the reference pieces built here)")


def pytest_collection_modifyitems(config, items):
    import oracle_lib
    if not oracle_lib.ref_available():
        skip = pytest.mark.skip(reason="oracle/_ref not built (reference tree absent)")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
