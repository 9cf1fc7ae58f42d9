import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "plonky3-gpu_amd"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and the built libp3hip.so")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def orc():
    import native

    return native.oracle()


@pytest.fixture(scope="session")
def ref():
    import native

    r = native.reference()
    if r is None:
        pytest.skip("reference build oracle/_ref/libp3ref.so not present")
    return r


@pytest.fixture(scope="session")
def ctx(golden):
    """A libp3hip context on device 0 with the fixture Poseidon2 constants.
    There is no CPU fallback: a missing library or device is an error."""
    import p3hip

    c = p3hip.Context(0)
    c.set_poseidon2(p3hip.BABYBEAR, golden["consts_bb"], 8, 13)
    c.set_poseidon2(p3hip.GOLDILOCKS, golden["consts_gl"], 8, 22)
    yield c
    c.close()
