import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    from dygrass_amd import Context

    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def grid_wl():
    from dygrass_amd import synth

    return synth.config_grid_insert(1.0, 0)


@pytest.fixture(scope="session")
def mesh_wl():
    from dygrass_amd import synth

    return synth.config_mesh_delete(0.004, 0)


def csr(E):
    from dygrass_amd import csr_from_edges

    return csr_from_edges(E.V, E.u, E.v, E.w)


def f32bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)
