"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the
CPU (oracle, host-side parsing, ABI exports)."""
import gzip
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X")


@pytest.fixture(scope="session")
def models_dir(tmp_path_factory):
    """The reference's OBJ fixtures (tests/golden/models/*.obj.gz), unpacked."""
    d = tmp_path_factory.mktemp("models")
    src = os.path.join(TESTS, "golden", "models")
    for f in os.listdir(src):
        if f.endswith(".obj.gz"):
            with gzip.open(os.path.join(src, f), "rb") as fi, open(d / f[:-3], "wb") as fo:
                shutil.copyfileobj(fi, fo)
    return d


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def cv():
    """The product library; built when missing (hipcc cross-compiles)."""
    import convector_amd

    try:
        convector_amd.lib()
    except ImportError:
        from convector_amd import build

        build.build()
        convector_amd.lib()
    return convector_amd
