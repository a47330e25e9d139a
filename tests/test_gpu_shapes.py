"""Launch shapes of the tunable kernels: every candidate gives bit-identical
runs (pairs, states, reports) on a batch with contact and friction; shapes
outside the candidates are refused; the autotuner's coordinate search picks
one candidate per shape and leaves it set."""
import numpy as np
import pytest

import shell_amd as S
from conftest import imperfect
from shell_amd import autotune

pytestmark = pytest.mark.gpu

H = 0.01


def say(*a):
    print(*a, flush=True)


def _batch(icosphere):
    X3, T3 = icosphere(3)
    X4, T4 = icosphere(4)
    meshes = [(imperfect(X3, 0.01 * H, seed=3), T3), (imperfect(X4, 0.01 * H, seed=4), T4)]
    kw = dict(bulk_modulus=0.05, volume_final_ratio=0.85, volume_ramp_time=8.0, volume_at_prediction=1, contact=1,
              mu_s=0.3, mu_d=0.3, cell_size_mult=1.0, cell_margin=0.1, rebuild_trigger=1, damping_alpha=0.05)
    return S.ShellBatch(meshes, [dict(kw, dt=0.1, dt_max=0.1), dict(kw, dt=0.05, dt_max=0.05)])


def _run(b, shapes):
    b.set_shapes(shapes)
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.apply_initial_conditions()
    reps = b.advance(300)
    return reps, b.get_state(S.ALL_SHELLS)[0], [b.pairs(s) for s in range(b.S)]


def test_every_candidate_gives_the_same_run(icosphere):
    b = _batch(icosphere)
    ref_reps, ref_u, ref_pairs = _run(b, None)
    assert b.get_shapes() == S.DEFAULT_SHAPES
    assert sum(r["n_rebuilds"] for r in ref_reps) > 2
    for dim, cands in autotune.CANDIDATES.items():
        for c in cands:
            shapes = dict(S.DEFAULT_SHAPES, **{dim: c})
            say("case shapes", shapes)
            reps, u, pairs = _run(b, shapes)
            assert np.array_equal(u, ref_u)
            for s in range(b.S):
                assert np.array_equal(pairs[s][0], ref_pairs[s][0]) and np.array_equal(pairs[s][1], ref_pairs[s][1])
                assert reps[s]["n_rebuilds"] == ref_reps[s]["n_rebuilds"]
                assert reps[s]["residual"] == ref_reps[s]["residual"]
    with pytest.raises(S.ShellError):
        b.set_shapes(dict(S.DEFAULT_SHAPES, node_waves=5))
    b.close()


def test_autotuner_picks_a_candidate_per_shape(icosphere):
    b = _batch(icosphere)
    shapes, rows = autotune.tune(b, steps=20, log=say)
    assert b.get_shapes() == shapes
    for dim, cands in autotune.CANDIDATES.items():
        assert shapes[dim] in cands
        mine = [r for r in rows if r[0] == dim]
        assert [r[1] for r in mine] == list(cands) and sum(r[6] for r in mine) == 1
        assert len({r[3] for r in mine}) == 1           # the same launches for every candidate
    b.close()
