import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_count():
    """HIP devices seen by the product library itself (hipGetDeviceCount
    through libshellhip), not by torch: the product never uses torch"""
    import shell_amd
    return shell_amd.lib().shx_device_count()


def pytest_collection_modifyitems(config, items):
    markexpr = config.getoption("-m") or ""
    wants_gpu = "gpu" in markexpr and "not gpu" not in markexpr
    if not any("gpu" in it.keywords for it in items):
        return
    try:
        n = gpu_count()
    except Exception as e:
        if wants_gpu:
            raise pytest.UsageError(f"-m gpu selected but libshellhip cannot be loaded: {e}")
        n = 0
    if n > 0:
        return
    if wants_gpu:
        raise pytest.UsageError("-m gpu selected but hipGetDeviceCount reports no device")
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
