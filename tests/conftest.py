import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("dg-bgk_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def have_ref(prec=64):
    return os.path.exists(os.path.join(ROOT, "oracle", "_ref", f"libref{prec}.so"))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle_py import lib
    return lib()
