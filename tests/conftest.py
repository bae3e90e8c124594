"""Shared pytest setup: import paths, the ``gpu`` marker, common fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "crystallographicfft.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def group_setup(sg, N):
    """get_ops -> find_optimal_shift -> (R, t) arrays of the shifted ops."""
    import cfft_hip as cf
    ops = cf.get_ops(sg, len(N), N)
    shift, ops_s = cf.find_optimal_shift(ops, N)
    R, t = cf.ops_arrays(ops_s)
    return ops, ops_s, shift, R, t


@pytest.fixture(scope="session")
def setup_group():
    cache = {}

    def get(sg, N):
        key = (sg, tuple(N))
        if key not in cache:
            cache[key] = group_setup(sg, N)
        return cache[key]
    return get
