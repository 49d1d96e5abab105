import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "louvain-community-detection_amd"))

from loader import load_package  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def lv():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def bcsstk16_text():
    with gzip.open(os.path.join(GOLDEN, "bcsstk16.mtx.gz"), "rt") as f:
        return f.read()


@pytest.fixture(scope="session")
def bcsstk16_path(tmp_path_factory, bcsstk16_text):
    p = tmp_path_factory.mktemp("mtx") / "bcsstk16.mtx"
    p.write_text(bcsstk16_text)
    return str(p)
