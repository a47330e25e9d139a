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


def plate(n, a=1.0):
    """the square [0, a]^2 of the x-z plane (y = 0) as an n x n grid of
    squares, each cut by the same diagonal (interior valence 6; corners with
    one or two triangles, other boundary vertices with three).  Not the x-y
    plane: the reference's element frame projects the z axis onto the
    tangent plane (reinit_mod, ShellSystem.C:3339-3424), which vanishes where
    the normal is the z axis."""
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    X = np.stack([i.ravel() * a / n, np.zeros(i.size), j.ravel() * a / n], axis=1)
    idx = lambda p, q: p * (n + 1) + q
    T = []
    for p in range(n):
        for q in range(n):
            T.append((idx(p, q), idx(p + 1, q), idx(p + 1, q + 1)))
            T.append((idx(p, q), idx(p + 1, q + 1), idx(p, q + 1)))
    return X, np.array(T, dtype=np.int32)


def tube(m, k, r=0.5, length=1.0):
    """an open cylinder of radius r along z: m nodes around, k rows of
    squares cut by the same diagonal, counter-clockwise seen from outside"""
    ang = 2 * np.pi * np.arange(m) / m
    X = np.array([(r * np.cos(t), r * np.sin(t), length * q / k) for q in range(k + 1) for t in ang])
    idx = lambda j, q: q * m + j % m
    T = []
    for q in range(k):
        for j in range(m):
            T.append((idx(j, q), idx(j + 1, q), idx(j + 1, q + 1)))
            T.append((idx(j, q), idx(j + 1, q + 1), idx(j, q + 1)))
    return X, np.array(T, dtype=np.int32)
