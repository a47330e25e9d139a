import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def icosphere():
    import oracle
    cache = {}

    def get(level):
        if level not in cache:
            cache[level] = oracle.icosphere(level)
        return cache[level]
    return get


def imperfect(X, amp, seed):
    """radial Gaussian perturbation of each vertex, standard deviation amp"""
    rng = np.random.default_rng(seed)
    n = X / np.linalg.norm(X, axis=1)[:, None]
    return X + n * rng.normal(size=(len(X), 1)) * amp
