"""Shared fixtures. Tests marked `gpu` run the HIP kernels through librsim.so's C ABI and need an MI355X; the rest run on
the CPU (the oracle against the golden vectors and the reference build, the host logic, the C ABI's exports)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rocketsim-cuda_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (runs the gfx950 kernels through librsim.so)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    if not os.path.exists(orc.ORACLE_LIB):
        orc.build(ref=False)
    return orc.Oracle()


@pytest.fixture(scope="session")
def rsim():
    import rsim
    if not os.path.exists(rsim.LIB_PATH):  # only when nothing is built yet (a fresh checkout)
        from rsim import build as rb
        rb.build()
    rsim.lib()
    return rsim


@pytest.fixture(scope="session")
def world(rsim):
    return rsim.build_world(0, 0)


@pytest.fixture(scope="session")
def device(rsim):
    dev = rsim.SimDevice(0)
    yield dev
    dev.close()


@pytest.fixture(scope="session")
def dev_world(device, world):
    return device.upload_static(world)


def first_diff(a, b):
    """(arena, word) of the first differing word of two state arrays, or None."""
    d = np.argwhere(np.asarray(a) != np.asarray(b))
    return None if d.size == 0 else tuple(int(x) for x in d[0])
