import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def ref_lib():
    """The reference solver modules compiled from their own sources
    (oracle/_ref/libimexref.so); skips where they are not built."""
    from oracle import refpy
    if not refpy.available("ieee"):
        pytest.skip("oracle/_ref/libimexref.so not built (make -C oracle ref)")
    # the reference pauses on a READ from unit 5 after a physics error: give
    # it an empty standard input so such a pause returns at once
    fd = os.open(os.devnull, os.O_RDONLY)
    os.dup2(fd, 0)
    os.close(fd)
    return refpy


def golden_path(name):
    return os.path.join(GOLDEN, name)
