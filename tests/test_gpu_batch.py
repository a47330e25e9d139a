"""Batched shells: every phase of a step is one launch over all shells of
the batch.  Each shell of a batch must evolve exactly as it does alone,
bit for bit, with contact too (the same arithmetic, the same per-shell
reduction orders, contact rows sorted and contact forces summed exactly in
fixed point), with pair lists identical to the CPU oracle's.  Graph replay of step attempts must equal launching them one by
one.
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * 1.0 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


def _shells(icosphere):
    X3, T3 = icosphere(3)
    X4, T4 = icosphere(4)
    meshes = [(imperfect(X3, 0.01 * H, seed=1), T3), (imperfect(X4, 0.01 * H, seed=2), T4),
              (imperfect(X3, 0.02 * H, seed=3), T3)]
    params = [
        dict(pressure_final=-1.5 * PCR, pressure_ramp_time=20.0, dt=0.1, damping_alpha=0.05),
        dict(E2=0.3, bulk_modulus=0.05, volume_final_ratio=0.95, volume_ramp_time=30.0, dt=0.05, dt_max=0.05,
             damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.2, mu_d=0.2, cell_size_mult=1.0),
        dict(E2=0.8, thickness=0.015, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=20.0, dt=0.1,
             dt_max=0.1, damping_alpha=0.05, adapt_dt_every=1, volume_at_prediction=1, contact=1, mu_s=0.4,
             mu_d=0.4, cell_size_mult=1.0),
    ]
    return meshes, params


def test_batch_equals_single_shells(icosphere):
    meshes, params = _shells(icosphere)
    say("case batch of", len(meshes), "shells vs each alone")
    b = S.ShellBatch(meshes, params)
    singles = [S.ShellSystem(X, T, **p) for (X, T), p in zip(meshes, params)]
    for s, (X, _) in enumerate(meshes):
        z = np.zeros(3 * len(X))
        b.set_state(s, z)
        singles[s].set_state(z)
        singles[s].apply_initial_conditions()
    b.apply_initial_conditions()
    for k in range(3):
        rb = b.advance(60)
        for s, g in enumerate(singles):
            rs = g.advance(60)
            ub, vb, _, tb = b.get_state(s)
            us, vs, _, ts = g.get_state()
            say("  chunk", k, "shell", s, "steps", rb[s]["n_steps"], rs["n_steps"], "rej", rb[s]["n_rejected"],
                rs["n_rejected"], "pairs", rb[s]["n_pairs"], rs["n_pairs"], "rel u", relerr(ub, us))
            assert rb[s]["n_steps"] == rs["n_steps"] and rb[s]["n_rejected"] == rs["n_rejected"]
            assert rb[s]["n_rebuilds"] == rs["n_rebuilds"] and rb[s]["n_pairs"] == rs["n_pairs"]
            assert tb == ts
            assert np.array_equal(ub, us) and np.array_equal(vb, vs)
            assert rb[s]["residual"] == rs["residual"]
            assert rb[s]["e_self"] == rs["e_self"] and rb[s]["e_fric"] == rs["e_fric"]
            if params[s].get("contact"):
                e1b, e2b = b.pairs(s)
                e1s, e2s = g.pairs()
                assert np.array_equal(e1b, e1s) and np.array_equal(e2b, e2s)
    for g in singles:
        g.close()
    b.close()


def test_batch_shell_against_oracle(icosphere):
    """the contact shell with adaptive dt inside a batch against the oracle:
    pair lists bit-exact, rebuilds and rejections identical"""
    meshes, params = _shells(icosphere)
    say("case batch shell 2 vs oracle")
    b = S.ShellBatch(meshes, params)
    X, T = meshes[2]
    o = oracle.Oracle(X, T, **params[2])
    for s, (Xs, _) in enumerate(meshes):
        b.set_state(s, np.zeros(3 * len(Xs)))
    o.set_state(np.zeros(3 * len(X)))
    b.apply_initial_conditions()
    o.init_dynamics()
    for k in range(3):
        rb = b.advance(80)[2]
        ro = o.step(80)
        ub = b.get_state(2)[0]
        uo = o.get_state()[0]
        e1b, e2b = b.pairs(2)
        e1o, e2o = o.pairs()
        say("  steps", rb["n_steps"], "rej", rb["n_rejected"], ro["n_rejected"], "rebuilds", rb["n_rebuilds"],
            ro["n_rebuilds"], "pairs", len(e1b), len(e1o), "rel u", relerr(ub, uo))
        assert rb["n_rebuilds"] == ro["n_rebuilds"] and rb["n_rejected"] == ro["n_rejected"]
        assert np.array_equal(e1b, e1o) and np.array_equal(e2b, e2o)
        assert relerr(ub, uo) < 1e-9
    b.close()


@pytest.mark.parametrize("contact", [False, True])
def test_graph_replay_equals_launches(icosphere, contact):
    meshes, params = _shells(icosphere)
    if contact:
        meshes, params = [meshes[2], meshes[1]], [params[2], params[1]]
    else:
        meshes, params = [meshes[0], meshes[0]], [params[0], dict(params[0], dt=0.07)]
    say("case graph replay vs launches")
    runs = []
    for graphs in (True, False):
        b = S.ShellBatch(meshes, params)
        b.set_graphs(graphs)
        for s, (X, _) in enumerate(meshes):
            b.set_state(s, np.zeros(3 * len(X)))
        b.apply_initial_conditions()
        reps = b.advance(150)
        runs.append(([b.get_state(s)[0] for s in range(len(meshes))], reps))
        b.close()
    for s in range(len(meshes)):
        assert np.array_equal(runs[0][0][s], runs[1][0][s])
        assert runs[0][1][s]["n_steps"] == runs[1][1][s]["n_steps"] == 150
        assert runs[0][1][s]["e_self"] == runs[1][1][s]["e_self"]
