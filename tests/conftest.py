import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "ref: needs oracle/_ref (the reference build)")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    import harness
    if not os.path.exists(harness.ORACLE_LIB) or (
            os.path.isdir(harness.REFERENCE_TREE) and not os.path.exists(harness.REF_LIB)):
        harness.build_oracle()
