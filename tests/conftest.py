import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def sp():
    import sparsep_amd.api as api
    api.lib()
    return api


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def ctx(sp):
    # No GPU node here: skip.  On a GPU box a failing context is an error.
    if not sp.gpu_present():
        pytest.skip("no GPU")
    return sp.Context(0)


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
