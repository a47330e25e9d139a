"""Host side of the learned warm start: parameter draws, the shells they
define, rigid-motion removal and the principal components (no GPU)."""
import numpy as np

from shell_amd import warmstart as W
from shell_amd.surrogate import pca
from shell_amd.workloads import icosphere


def _rot(ax, th):
    ax = np.asarray(ax, float) / np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def test_draw_params_ranges():
    P = W.draw_params(4000, seed=3)
    assert P.shape == (4000, 6)
    lo = [0.3, 0.25, 0.008, 0.86, 0.0, 0.005]
    hi = [1.0, 0.35, 0.016, 0.92, 0.5, 0.02]
    assert np.all(P >= lo) and np.all(P <= hi)
    # log-uniform thickness: log h is uniform, so its mean is the log midpoint
    assert abs(np.mean(np.log(P[:, 2])) - 0.5 * (np.log(0.008) + np.log(0.016))) < 0.01
    assert np.array_equal(P, W.draw_params(4000, seed=3))


def test_shells_share_one_imperfection_pattern():
    P = W.draw_params(3, seed=1)
    meshes, prms = W.shells(P)
    Xb, _ = icosphere(W.LEVEL)
    n = Xb / np.linalg.norm(Xb, axis=1)[:, None]
    dev = [np.einsum("ij,ij->i", X - Xb, n) for X, _ in meshes]
    # radial deviation / (amplitude * h) is the same pattern for every shell
    pat = [d / (p[5] * p[2]) for d, p in zip(dev, P)]
    assert np.allclose(pat[0], pat[1], atol=1e-9) and np.allclose(pat[0], pat[2], atol=1e-9)
    assert abs(np.std(pat[0]) - 1.0) < 0.05
    for p, q in zip(prms, P):
        assert p["E2"] == q[0] and p["nu"] == q[1] and p["thickness"] == q[2]
        assert p["volume_final_ratio"] == q[3] and p["mu_s"] == p["mu_d"] == q[4]
        assert p["volume_ramp_time"] == W.RAMP_TIME


def test_align_removes_rigid_motion_only():
    rng = np.random.default_rng(0)
    X = rng.normal(size=(200, 3))
    u = 1e-3 * rng.normal(size=X.shape)
    a0 = W.align([X], u.reshape(1, -1))[0].reshape(-1, 3)
    R, c = _rot([1, 2, 3], 0.7), np.array([0.3, -1.0, 2.0])
    x = (X + u) @ R.T + c
    a1 = W.align([X], (x - X).reshape(1, -1))[0].reshape(-1, 3)
    assert np.abs(a1 - a0).max() < 1e-12
    # a pure rigid motion aligns to zero
    xr = X @ R.T + c
    assert np.abs(W.align([X], (xr - X).reshape(1, -1))).max() < 1e-12


def test_pca_reconstructs_low_rank_data():
    rng = np.random.default_rng(1)
    B = rng.normal(size=(5, 300))
    C = rng.normal(size=(40, 5))
    U = C @ B + 2.0
    umean, basis, coef, kept = pca(U, 5)
    assert basis.shape == (5, 300) and coef.shape == (40, 5)
    assert np.allclose(basis @ basis.T, np.eye(5), atol=1e-12)
    assert np.allclose(umean + coef @ basis, U, atol=1e-10)
    assert abs(kept - 1.0) < 1e-12
    _, _, _, kept2 = pca(U, 2)
    assert kept2 < 1.0
