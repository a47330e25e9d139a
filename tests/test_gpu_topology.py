"""The topology half of build_cache on the GPU (shx_topo.hip) against the
CPU oracle: rotated triangles, valences, one-ring patches and contact
exclusion sets bit for bit, for single shells and a batch of mixed levels;
and the mesh errors the preprocessor reports (ShellSystem.C:490-571,
657-741; the reference's subdivision mesh requirements)."""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("level", [2, 3, 5])
def test_patches_equal_the_oracle(icosphere, level):
    say("case patches level", level)
    X, T = icosphere(level)
    X = imperfect(X, 1e-4, seed=level)
    g = S.ShellSystem(X, T)
    o = oracle.Oracle(X, T)
    gi, gt = g.patches()
    oi, ot, _ = o.patches()
    assert np.array_equal(gi, oi) and np.array_equal(gt, ot)
    g.close()
    o.close()


@pytest.mark.parametrize("level,h,margin", [(2, 0.4, 0.4), (3, 0.05, 0.4), (4, 0.01, 0.4)])
def test_close_sets_equal_the_oracle(icosphere, level, h, margin):
    """the coarse case's cutoff reaches several rings out, so the search
    iterates; the others are the ring of vertex neighbours and a few more"""
    say("case close sets level", level, "h", h)
    X, T = icosphere(level)
    X = imperfect(X, 1e-3 * h, seed=7 + level)
    kw = dict(thickness=h, neighbor_margin=margin, contact=1)
    g = S.ShellSystem(X, T, **kw)
    o = oracle.Oracle(X, T, **kw)
    gs, gids = g.close_sets()
    os_, oids = o.close_sets()
    say("  ids", len(gids), len(oids), "max set", np.diff(gs).max())
    assert np.array_equal(gs, os_) and np.array_equal(gids, oids)
    g.close()
    o.close()


def test_batch_of_mixed_levels(icosphere):
    """one preprocessor launch per phase for all shells: each shell's patches
    and exclusion sets equal the oracle's"""
    say("case batch topology")
    meshes, params = [], []
    for k, level in enumerate((3, 4, 2, 3)):
        X, T = icosphere(level)
        meshes.append((imperfect(X, 1e-4, seed=20 + k), T))
        params.append(dict(thickness=0.01 * (k + 1), contact=1 if k != 2 else 0))
    b = S.ShellBatch(meshes, params)
    for s, ((X, T), p) in enumerate(zip(meshes, params)):
        o = oracle.Oracle(X, T, **p)
        gi, gt = b.shell_patches(s)
        oi, ot, _ = o.patches()
        assert np.array_equal(gi, oi) and np.array_equal(gt, ot)
        gs, gids = b.close_sets(s)
        if p["contact"]:
            os_, oids = o.close_sets()
            assert np.array_equal(gs, os_) and np.array_equal(gids, oids)
        else:
            assert len(gids) == 0
        o.close()
    b.close()


def test_mesh_errors(icosphere):
    X, T = icosphere(3)
    with pytest.raises(S.ShellError, match="not closed"):
        S.ShellSystem(X, T[1:])
    T2 = T.copy()
    T2[5] = T2[5, ::-1]
    with pytest.raises(S.ShellError, match="non-manifold or inconsistently oriented"):
        S.ShellSystem(X, T2)
    X0, T0 = icosphere(0)            # the icosahedron: every vertex has valence 5
    with pytest.raises(S.ShellError, match="more than one irregular vertex"):
        S.ShellSystem(X0, T0)
    T3 = T.copy()
    T3[0, 0] = len(X)
    with pytest.raises(S.ShellError, match="vertex index out of range"):
        S.ShellSystem(X, T3)
