import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on device 0)")


@pytest.fixture(scope="session")
def ctx():
    from anukulan_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ref():
    import _oracle
    if _oracle.ref is None:
        pytest.skip("oracle/_ref/libanukulan_ref.so not built (needs the reference sources)")
    return _oracle


@pytest.fixture(scope="session")
def orc():
    import _oracle
    if _oracle.oracle is None:
        pytest.fail("oracle/liboracle.so not built: run `make -C oracle oracle`")
    return _oracle
