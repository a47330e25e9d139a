"""GPU (gfx950) path against the CPU FP64 oracle on the same inputs.

Bars: contact pair lists bit-exact; tables, masses, forces and trajectories
within a relative tolerance of 1e-10 (FP64 path; the summation order of the
node gather differs from the oracle's serial element-order assembly).
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * 1.0 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))
RTOL = 1e-10


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


def pair(X, T, **kw):
    return S.ShellSystem(X, T, **kw), oracle.Oracle(X, T, **kw)


@pytest.mark.parametrize("level", [3, 4])
def test_tables_and_mass(icosphere, level):
    say("case tables level", level)
    X, T = icosphere(level)
    X = imperfect(X, 0.01 * H, seed=level)
    g, o = pair(X, T)
    gi, gt = g.patches()
    oi, ot, _ = o.patches()
    assert np.array_equal(gi, oi) and np.array_equal(gt, ot)
    tg, to = g.elem_tables(), o.elem_tables()
    # reinit_mod scales the east tangent by |b1| sin(angle(b1, north)): when
    # an element's first tangent points almost north the frame is
    # ill-conditioned and rounding differences grow like 1/sin^2 (the
    # reference also forms an extra Rodrigues rotation that cancels
    # mathematically and that the GPU does not form)
    assert relerr(tg[:, :72], to[:, :72]) < 1e-8
    assert relerr(tg[:, 72:101], to[:, 72:101]) < 1e-8
    mg, cg = g.mass()
    mo, co = o.mass()
    assert relerr(mg, mo) < 1e-13 and relerr(cg, co) < 1e-13
    g.close()


@pytest.mark.parametrize("level", [3, 4])
def test_forces_random_displacement(icosphere, level):
    say("case forces level", level)
    X, T = icosphere(level)
    X = imperfect(X, 0.01 * H, seed=10 + level)
    kw = dict(pressure_final=-2 * PCR)
    g, o = pair(X, T, **kw)
    rng = np.random.default_rng(level)
    for scale in (1e-5, 1e-3, 3e-2):
        u = rng.normal(size=3 * len(X)) * scale
        fg, feg, eng, ag = g.forces(u, t=0.5)
        fo, feo, eno, ao = o.forces(u, t=0.5)
        say("  scale", scale, "rel f", relerr(fg, fo), "rel fext", relerr(feg, feo))
        assert relerr(fg, fo) < RTOL
        assert relerr(feg, feo) < RTOL
        assert relerr(ag, ao) < 1e-13
        assert abs(eng[0] - eno[0]) <= 1e-10 * abs(eno[0])
        assert abs(eng[1] - eno[1]) <= 1e-10 * abs(eno[1])
    g.close()


def test_trajectory_constant_dt(icosphere):
    say("case trajectory L3 constant dt")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=3)
    kw = dict(pressure_final=-2 * PCR, pressure_ramp_time=5.0, dt=0.1, damping_alpha=0.05)
    g, o = pair(X, T, **kw)
    rng = np.random.default_rng(5)
    u0 = rng.normal(size=3 * len(X)) * 1e-4
    g.set_state(u0)
    o.set_state(u0)
    g.apply_initial_conditions()
    o.init_dynamics()
    for k in range(4):
        rg = g.advance(50)
        ro = o.step(50)
        ug, vg, ag, tg = g.get_state()
        uo, vo, ao, to = o.get_state()
        say("  steps", rg["n_steps"], "t", tg, to, "rel u", relerr(ug, uo), "rel v", relerr(vg, vo),
            "res", rg["residual"], ro["residual"])
        assert abs(tg - to) < 1e-12
        assert relerr(ug, uo) < RTOL and relerr(vg, vo) < RTOL and relerr(ag, ao) < 1e-9
        assert abs(rg["residual"] - ro["residual"]) <= 1e-8 * ro["residual"]
    g.close()


def _pancake(X, s):
    Y = X.copy()
    Y[:, 2] *= s
    return (Y - X).reshape(-1)


@pytest.mark.parametrize("level,squash", [(3, 0.006), (4, 0.008)])
def test_contact_pairs_bitexact(icosphere, level, squash):
    """pair lists of a pancaked shell bit-exact, contact and total force to
    RTOL; its 256-element tiles hold up to 2173 (L3) and 3345 (L4) pairs, so
    every thread of the contact kernel walks a run of pairs across rows"""
    say("case contact pairs level", level)
    X, T = icosphere(level)
    kw = dict(contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
    g, o = pair(X, T, **kw)
    u = _pancake(X, squash)
    g.set_state(u)
    o.set_state(u)
    ng = g.compute_shell_shell_contact_candidates(force=True)
    no = o.rebuild(force=True)
    e1g, e2g = g.pairs()
    e1o, e2o = o.pairs()
    say("  pairs", ng, no)
    assert ng == no and ng > 0
    assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o)
    fg, feg, eng, _ = g.forces(u, t=0.0)
    fo, feo, eno, _ = o.forces(u, t=0.0)
    say("  contact force rel", relerr(feg, feo), "Eself", eng[2], eno[2])
    assert relerr(feg, feo) < RTOL
    assert relerr(fg, fo) < RTOL
    g.close()


@pytest.mark.parametrize("trigger", [0, 1])
def test_contact_dynamics_adaptive(icosphere, trigger):
    """Squashed L3 shell whose two sheets close at 0.04 per unit time, with
    self-contact, friction and the reference's adaptive dt: rebuild counts,
    rejected attempts and sorted pair lists identical after every chunk,
    states within tolerance over the window; with the reference's rebuild
    trigger and with the displacement trigger (extension)."""
    say("case contact dynamics L3, rebuild trigger", trigger)
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=7)
    kw = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1, contact=1, mu_s=0.3, mu_d=0.3,
              cell_size_mult=1.0, rebuild_trigger=trigger)
    g, o = pair(X, T, **kw)
    u = _pancake(X, 0.03)
    v = np.zeros_like(u).reshape(-1, 3)
    v[:, 2] = -0.02 * np.sign(X[:, 2])
    v = v.reshape(-1)
    g.set_state(u, v)
    o.set_state(u, v)
    g.apply_initial_conditions()
    o.init_dynamics()
    seen_pairs = 0
    for k in range(8):
        rg = g.advance(25)
        ro = o.step(25)
        ug, vg, _, tg = g.get_state()
        uo, vo, _, to = o.get_state()
        e1g, e2g = g.pairs()
        e1o, e2o = o.pairs()
        seen_pairs = max(seen_pairs, len(e1o))
        say("  steps", rg["n_steps"], "t", tg, to, "rej", rg["n_rejected"], ro["n_rejected"], "pairs", len(e1g),
            len(e1o), "contacts", rg["n_contacts"], ro["n_contacts"], "rebuilds", rg["n_rebuilds"],
            ro["n_rebuilds"], "rel u", relerr(ug, uo))
        assert rg["n_rebuilds"] == ro["n_rebuilds"] and rg["n_rejected"] == ro["n_rejected"]
        assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o)
        assert abs(tg - to) <= 1e-12 * max(1, abs(to))
        assert relerr(ug, uo) < 1e-9
    assert seen_pairs > 0
    g.close()


def test_deflation_trajectory_and_residual(icosphere):
    """volume-controlled deflation (bulk modulus, volume of the force
    configuration), post-buckling: states, rigid-free residual and the
    convergence step agree with the oracle"""
    say("case deflation L3")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=2024)
    kw = dict(bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=60.0, dt=0.1175, dt_max=0.1175,
              damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
    g, o = pair(X, T, **kw)
    z = np.zeros(3 * len(X))
    g.set_state(z)
    o.set_state(z)
    g.apply_initial_conditions()
    o.init_dynamics()
    for k in range(4):
        rg = g.advance(150)
        ro = o.step(150)
        ug = g.get_state()[0]
        uo = o.get_state()[0]
        say("  steps", rg["n_steps"], "rel u", relerr(ug, uo), "res", rg["residual"], ro["residual"], "resdef",
            rg["residual_deform"], ro["residual_deform"], "V", rg["volume"], ro["volume"])
        assert relerr(ug, uo) < 1e-9
        assert abs(rg["residual_deform"] - ro["residual_deform"]) <= 1e-6 * ro["residual_deform"] + 1e-12
        assert abs(rg["volume"] - ro["volume"]) <= 1e-12 * ro["volume"]
    g.close()
