"""Smooth (limit-surface) shell-shell contact on the GPU against the CPU FP64
oracle: SmoothShellShellContact::compute for single pairs (SmoothShellShell
Contact.C:12-204), the candidate search with the longer candidate distance
(ShellSystem.C:1940-1944, 2071-2101), the contact forces
(add_shell_shell_contact_force, :1792-1907, with the limit-surface closest
points), adaptive dynamics with friction, batches mixing smooth and linear
contact, and state-file restarts.

Bars: pair lists bit-exact; a batch equal to its shells run alone bit for
bit.  The closest points come out of an iterative search, so the two sides
agree to its stopping point: parameters within 1e-9 and distances within
1e-12 (relative) where the problem is well conditioned.  Two places are
not: within xi + eta < 1e-3 of an irregular vertex, where the limit
surface's parametrisation degenerates (tangents vanish like (2 lambda)^n,
the 4^n-scaled second derivatives are rounding; SubdivEvaluation's own
method), the parameters agree within 1e-4 and d2 within 1e-8; and pairs of
distant elements can have several local minima exactly tied on a
symmetric mesh, of which each side may find another (d2 equal, parameters
not).  Hence forces within 1e-6 and trajectories within 1e-6 (relative).
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


def pancake(X, s):
    Y = X.copy()
    Y[:, 2] *= s
    return (Y - X).reshape(-1)


KW = dict(contact=2, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
D2_FRICTION = (1.1 * H) ** 2       # (1 + static_friction_tol) h, squared


def near_irregular(x):
    """the pair's parameters lie within xi + eta < 1e-3 of an element's node
    0 (the irregular vertex of an irregular element)"""
    return x[0] + x[1] < 1e-3 or x[2] + x[3] < 1e-3


def compare_parameters(vg, vo, d2g, d2o):
    """pairs in contact range: parameters within 1e-9 (1e-4 near an
    irregular vertex); every pair: d2 within 1e-8 relative"""
    worst = 0.0
    for a, b, dg, do in zip(vg, vo, d2g, d2o):
        assert abs(dg - do) <= 1e-8 * do
        if do < D2_FRICTION:
            diff = np.abs(a - b).max()
            assert diff < (1e-4 if near_irregular(b) else 1e-9), (a, b)
            if not near_irregular(b):
                worst = max(worst, diff)
    return worst


def test_single_pairs_equal_the_oracle(icosphere):
    """the entry point for one pair, from the centroids and from a given
    start, at a pancaked state: minimiser, distance and contact point"""
    say("case smooth single pairs")
    X, T = icosphere(3)
    X = imperfect(X, 1e-4, seed=3)
    g = S.ShellSystem(X, T, **KW)
    o = oracle.Oracle(X, T, **KW)
    u = pancake(X, 0.08)
    g.set_state(u)
    o.set_state(u)
    Xs = X + u.reshape(-1, 3)
    _, tri = g.patches()
    cent = Xs[tri].mean(axis=1)
    top, bot = np.where(cent[:, 2] > 0.03)[0], np.where(cent[:, 2] < -0.03)[0]
    rng = np.random.default_rng(8)
    worst = 0.0
    for k in range(200):
        e1 = int(rng.choice(top))
        e2 = int(bot[np.argsort(np.linalg.norm(cent[bot, :2] - cent[e1, :2], axis=1))[rng.integers(0, 8)]])
        x0 = (1 / 3,) * 4 if k % 2 == 0 else tuple(rng.dirichlet((1, 1, 1))[:2]) + tuple(rng.dirichlet((1, 1, 1))[:2])
        a = S.smooth_contact_compute(g, e1, e2, x0)
        b = o.smooth_contact(e1, e2, x0)
        if near_irregular(b["x"]):
            assert np.abs(a["x"] - b["x"]).max() < 1e-4 and abs(a["d2"] - b["d2"]) <= 1e-8 * b["d2"]
            continue
        worst = max(worst, np.abs(a["x"] - b["x"]).max())
        assert np.abs(a["x"] - b["x"]).max() < 1e-9, (e1, e2)
        assert abs(a["d2"] - b["d2"]) <= 1e-12 * b["d2"]
        assert np.abs(a["point"] - b["point"]).max() < 1e-12 and a["problem"] == 0
    say("  worst parameter difference", worst)
    g.close()
    o.close()


@pytest.mark.parametrize("squash", [0.006, 0.012])
def test_pairs_and_contact_forces(icosphere, squash):
    """a pancaked L3 shell: about 41 000 candidates (up to 82 in one row:
    the candidate distance is an element size longer than linear contact's),
    the lists bit-exact, each pair's closest-point parameters and the
    contact force equal to the oracle's"""
    say("case smooth pairs and forces, squash", squash)
    X, T = icosphere(3)
    # the icosphere's mirror symmetries give exactly tied search paths (each
    # side may then take another basin); a 1e-4 imperfection breaks them
    X = imperfect(X, 1e-4, seed=5)
    g = S.ShellSystem(X, T, **KW)
    o = oracle.Oracle(X, T, **KW)
    u = pancake(X, squash)
    g.set_state(u)
    o.set_state(u)
    ng = g.compute_shell_shell_contact_candidates(force=True)
    no = o.rebuild(force=True)
    e1g, e2g = g.pairs()
    e1o, e2o = o.pairs()
    say("  pairs", ng, no, "max row", np.bincount(e1o).max())
    assert ng == no and ng > 0
    assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o)
    sg = g.pair_state()
    vo = o.pair_coords()
    d2o = o.pair_state()[0]
    say("  parameters max diff", np.abs(sg["vw"] - vo).max(), "d2 rel", np.abs(sg["d2"] - d2o).max() / d2o.max(),
        "in contact range", int((d2o < D2_FRICTION).sum()))
    worst = compare_parameters(sg["vw"], vo, sg["d2"], d2o)
    say("  well-conditioned parameters max diff", worst)
    fg, feg, eng, _ = g.forces(u, t=0.0)
    fo, feo, eno, _ = o.forces(u, t=0.0)
    say("  contact force rel", relerr(feg, feo), "Eself", eng[2], eno[2])
    assert eno[2] > 0
    assert relerr(feg, feo) < 1e-6 and relerr(fg, fo) < 1e-6
    assert abs(eng[2] - eno[2]) <= 1e-8 * eno[2]
    g.close()
    o.close()


def test_smooth_contact_dynamics(icosphere):
    """the squashed L3 shell whose sheets close at 0.04 per unit time, with
    friction and adaptive dt: rebuild counts, rejected attempts and pair
    lists identical, states and the pairs' parameters within tolerance"""
    say("case smooth contact dynamics")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=7)
    kw = dict(KW, dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1)
    g = S.ShellSystem(X, T, **kw)
    o = oracle.Oracle(X, T, **kw)
    u = pancake(X, 0.03)
    v = np.zeros_like(u).reshape(-1, 3)
    v[:, 2] = -0.02 * np.sign(X[:, 2])
    v = v.reshape(-1)
    g.set_state(u, v)
    o.set_state(u, v)
    g.apply_initial_conditions()
    o.init_dynamics()
    for k in range(2):
        rg = g.advance(25)
        ro = o.step(25)
        ug, vg, _, tg = g.get_state()
        uo, vo, _, to = o.get_state()
        e1g, e2g = g.pairs()
        e1o, e2o = o.pairs()
        say("  t", tg, to, "rej", rg["n_rejected"], ro["n_rejected"], "rebuilds", rg["n_rebuilds"], ro["n_rebuilds"],
            "pairs", len(e1g), len(e1o), "rel u", relerr(ug, uo))
        assert rg["n_rebuilds"] == ro["n_rebuilds"] and rg["n_rejected"] == ro["n_rejected"]
        assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o)
        assert abs(tg - to) <= 1e-12 * max(1, abs(to))
        assert relerr(ug, uo) < 1e-6 and relerr(vg, vo) < 1e-5
        sg = g.pair_state()
        d2o = o.pair_state()[0]
        inr = d2o < D2_FRICTION
        say("  pairs in contact range", int(inr.sum()), "parameters max diff there",
            np.abs(sg["vw"][inr] - o.pair_coords()[inr]).max() if inr.any() else 0)
    g.close()
    o.close()


def test_batch_mixing_smooth_and_linear_equals_standalone(icosphere):
    """one batch holding a smooth-contact and a linear-contact shell (both
    through the smooth kernels and 256-slot rows) against each run alone
    (the linear one with the plain kernels and 64-slot rows): bit for bit"""
    say("case smooth and linear batch")
    X, T = icosphere(3)
    meshes = [(imperfect(X, 0.01 * H, seed=31 + k), T) for k in range(2)]
    base = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
    params = [dict(base, contact=2), dict(base, contact=1)]
    b = S.ShellBatch(meshes, params)
    states = []
    for (Xk, _), p in zip(meshes, params):
        u = pancake(Xk, 0.03)
        v = np.zeros_like(u).reshape(-1, 3)
        v[:, 2] = -0.02 * np.sign(Xk[:, 2])
        states.append((u, v.reshape(-1)))
    for s, (u, v) in enumerate(states):
        b.set_state(s, u, v=v, a=np.zeros_like(u), t=0.0)
    b.apply_initial_conditions()
    b.advance(40)
    for s, ((Xk, Tk), p) in enumerate(zip(meshes, params)):
        g = S.ShellSystem(Xk, Tk, **p)
        g.set_state(*states[s])
        g.apply_initial_conditions()
        g.advance(40)
        ug, vg, _, tg = g.get_state()
        ub, vb, _, tb = b.get_state(s)
        say("  shell", s, "t", tb, tg, "max |du|", np.abs(ub - ug).max())
        assert tb == tg and np.array_equal(ub, ug) and np.array_equal(vb, vg)
        assert np.array_equal(np.concatenate(b.pairs(s)), np.concatenate(g.pairs()))
        g.close()
    b.close()


def test_state_file_restart_with_smooth_contact(icosphere, tmp_path):
    """the sticking contacts travel in the state file (ShellSystem.C:
    3003-3025, 3093-3121) and restart their search at the centroids, as the
    reference's loaded SmoothShellShellContact does (:3116-3119); the
    continued run equals the uninterrupted one to the search's rounding
    (adaptive dt: t within 1e-8)"""
    say("case smooth restart")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=9)
    p = dict(KW, dt=0.01, dt_max=0.02, adapt_dt_every=1, damping_alpha=0.05, mu_s=0.6, mu_d=0.5,
             friction_stop_velocity=0.2)
    u = pancake(X, 0.012)
    v = np.zeros_like(u).reshape(-1, 3)
    v[:, 2] = -0.005 * np.sign(X[:, 2])
    v[:, 0] = 0.01
    a = S.ShellSystem(X, T, **p)
    a.set_state(u, v.reshape(-1))
    a.apply_initial_conditions()
    r1 = a.advance(20)
    n_stick = int(a.pair_state()["stick"].sum())
    path = str(tmp_path / "state.dat")
    a.save_state(path, 20)
    r2 = a.advance(20)
    ua, va, _, ta = a.get_state()
    b = S.ShellSystem(X, T, **p)
    b.load_state(path)
    r3 = b.advance(20)
    ub, vb, _, tb = b.get_state()
    say("  sticking contacts saved", n_stick, "contacts", r1["n_contacts"], "t", ta, tb, "rel u", relerr(ub, ua),
        "rel v", relerr(vb, va), "problems", r1["problem"], r2["problem"], r3["problem"])
    assert n_stick > 0 and r1["problem"] == r2["problem"] == r3["problem"] == 0
    # the restarted searches differ from the warm ones by rounding, which the
    # adaptive step size (from the local error estimate) carries into t
    assert abs(ta - tb) <= 1e-8 * ta and relerr(ub, ua) < 1e-6 and relerr(vb, va) < 1e-5
    a.close()
    b.close()
