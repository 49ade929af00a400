import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle.py_oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.py_oracle import Reference, reference_available

    if not reference_available():
        pytest.skip("oracle/_ref not built (needs the reference tree at build time)")
    return Reference()


GOLDEN = os.path.join(ROOT, "tests", "golden")
