import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
# tests tune launch shapes in memory only (no cache file read or written)
os.environ.setdefault('QSV_TUNE_CACHE', '')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU and the built libqsv.so')


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def random_real_pv(shape, rng):
    return rng.standard_normal(shape)
