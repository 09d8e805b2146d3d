import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def skeletal():
    from multifil_jax_amd.params import get_skeletal_params
    return get_skeletal_params()


@pytest.fixture(scope="session")
def topo22(skeletal):
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    return SarcTopology.create(2, 2, s, d)


@pytest.fixture(scope="session")
def topo44(skeletal):
    from multifil_jax_amd.topology import SarcTopology
    s, d = skeletal
    return SarcTopology.create(4, 4, s, d)


@pytest.fixture(scope="session")
def pdict(skeletal):
    from multifil_jax_amd.params import DYNAMIC_FIELDS
    _, d = skeletal
    return {n: float(getattr(d, n)) for n in DYNAMIC_FIELDS}


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
