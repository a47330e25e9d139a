"""The loads of compute_external_force after the pressure on the GPU against
the CPU FP64 oracle (ShellSystem.C:1587-1786): gravity, cavity contact with
its friction springs and normal damping (box, cylinder, ellipsoid; semi-axes
as expressions in t), and the rotation damping force.

Bars: forces and trajectories within 1e-10 relative (FP64 path; the node
gather sums in another order than the oracle's serial assembly), contact
counts and the state of the cavity springs equal, a batch with cavity shells
equal to the shells run alone bit for bit.
"""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01
RTOL = 1e-10


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


def split(R0=(1.0, 1.0, 1.0), rate=0.0, **kw):
    """the same configuration for the library (semi-axes as expressions in
    t when rate != 0) and for the oracle (R(t) = R0 exp(rate t))"""
    g = dict(kw)
    o = dict(kw)
    if kw.get("cavity_contact"):
        g["cavity_R"] = R0
        if rate:
            g["cavity_fn"] = tuple(f"{r!r}*exp({rate!r}*t)" for r in R0)
        o["cavity_R0"] = R0
        o["cavity_rate"] = rate
    return g, o


CASES = {
    "ellipsoid, friction, damping": dict(cavity_contact=1, cavity=2, R0=(1.003, 0.99, 0.985), mu_s_cavity=0.3,
                                         mu_d_cavity=0.2, cavity_damping=0.4, cavity_friction_damping=0.1),
    "box, dynamic friction": dict(cavity_contact=1, cavity=0, R0=(1.0, 1.2, 1.02), mu_d_cavity=0.25),
    "cylinder, static friction": dict(cavity_contact=1, cavity=1, R0=(1.01, 1.0, 0.99), mu_s_cavity=0.4,
                                      mu_d_cavity=0.3),
    "gravity and rotation damping": dict(gravity=0.02, gravity_dir=0, damping_rotation=0.3),
    "all of them, shrinking cavity": dict(cavity_contact=1, cavity=2, R0=(1.0, 1.0, 1.0), rate=-0.5, mu_s_cavity=0.2,
                                          mu_d_cavity=0.15, cavity_damping=0.2, gravity=0.01, gravity_dir=2,
                                          damping_rotation=0.1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_forces_with_body_loads(icosphere, case):
    say("case forces", case)
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=21)
    kg, ko = split(**CASES[case])
    g = S.ShellSystem(X, T, **kg)
    o = oracle.Oracle(X, T, **ko)
    rng = np.random.default_rng(7)
    for k, scale in enumerate((1e-4, 3e-3)):
        u = rng.normal(size=X.size) * scale
        v = rng.normal(size=X.size) * (0.2 if k else 3e-4)    # slow (sticking) and fast (slipping) nodes
        fg, feg, _, _ = g.forces(u, v, t=0.2)
        fo, feo, _, _ = o.forces(u, v, t=0.2)
        say("  scale", scale, "rel f", relerr(fg, fo), "rel fext", relerr(feg, feo), "|fext|", np.linalg.norm(feo))
        assert np.linalg.norm(feo) > 0
        assert relerr(fg, fo) < RTOL and relerr(feg, feo) < RTOL
        if CASES[case].get("cavity_contact"):
            so_g = g.cavity_springs()
            so_o = o.node_cavity()[0]
            stick_g, stick_o = np.any(so_g != 0, axis=1), np.any(so_o != 0, axis=1)
            say("  springs", stick_g.sum(), stick_o.sum())
            assert np.array_equal(stick_g, stick_o)
            assert relerr(so_g, so_o) < 1e-13
    g.close()
    o.close()


@pytest.mark.parametrize("case", ["ellipsoid, friction, damping", "all of them, shrinking cavity"])
def test_trajectory_with_body_loads(icosphere, case):
    """adaptive dt (the contact velocity clamp applies with cavity contact,
    ShellSystem.C:1037-1043), 200 steps, state, time, cavity energy and
    count every 50 steps.  The cavity squeezes the shell far beyond its
    buckling load, so rounding differences grow about e^28 per time unit
    (a physical instability: 1e-15 at t = 0.18, 2e-10 at t = 0.54 over 400
    steps); the states are compared while that growth stays below the bar"""
    say("case trajectory", case)
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=22)
    base = dict(dt=0.02, dt_max=0.05, adapt_dt_every=1, damping_alpha=0.05)
    kg, ko = split(**CASES[case], **base)
    g = S.ShellSystem(X, T, **kg)
    o = oracle.Oracle(X, T, **ko)
    z = np.zeros(X.size)
    g.set_state(z)
    o.set_state(z)
    g.apply_initial_conditions()
    o.init_dynamics()
    for k in range(4):
        rg = g.advance(50)
        ro = o.step(50)
        ug, vg, _, tg = g.get_state()
        uo, vo, _, to = o.get_state()
        say("  t", tg, to, "rel u", relerr(ug, uo), "rel v", relerr(vg, vo), "e_cav", rg["e_cav"], ro["e_cav"],
            "n_cav", rg["n_cav_contacts"], ro["n_cav"], "rejected", rg["n_rejected"], ro["n_rejected"])
        assert abs(tg - to) <= 1e-12 * abs(to)
        assert rg["n_rejected"] == ro["n_rejected"]
        assert relerr(ug, uo) < RTOL and relerr(vg, vo) < 1e-9
        assert rg["n_cav_contacts"] == ro["n_cav"]
        assert abs(rg["e_cav"] - ro["e_cav"]) <= 1e-9 * max(abs(ro["e_cav"]), 1e-300)
    if CASES[case].get("cavity_contact"):
        assert ro["n_cav"] > 0
    g.close()
    o.close()


def test_cavity_measurements(icosphere):
    """print_properties with cavity contact (density, cavity pressure,
    Nc_cav, E_cav; ShellSystem.C:2818-2898) and the ecav / efric node fields
    (ShellSystem.C:1626-1706) equal the oracle's"""
    say("case cavity measurements")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=23)
    kg, ko = split(**CASES["all of them, shrinking cavity"], dt=0.004, adapt_dt_every=0)
    g = S.ShellSystem(X, T, **kg)
    o = oracle.Oracle(X, T, **ko)
    z = np.zeros(X.size)
    g.set_state(z)
    o.set_state(z)
    g.apply_initial_conditions()
    o.init_dynamics()
    g.measure()
    o.measure()
    g.advance(30)
    o.step(30)
    pg = g.properties()
    po = o.properties()
    for k, name in enumerate(S.PROPERTY_COLUMNS):
        say("  ", name, pg[name], po[k])
        assert abs(pg[name] - po[k]) <= 1e-9 * max(abs(po[k]), 1e-12), name
    assert po[6] > 0 and po[11] > 0 and po[4] > 0
    u, v, _, t = g.get_state()
    nf, _ = g.fields()
    o.forces(u, v, t)
    _, ecav, efric = o.node_cavity()
    say("  fields ecav", relerr(nf[:, 0], ecav), "efric", relerr(nf[:, 1], efric))
    assert relerr(nf[:, 0], ecav) < 1e-9 and relerr(nf[:, 1], efric) < 1e-9
    # the fields evaluation leaves the run unchanged: the springs are those of the run
    g.close()
    o.close()


def test_batch_with_cavity_shells_equals_standalone(icosphere):
    """cavity, gravity and rotation damping shells next to a plain shell in
    one batch: each equals its standalone run bit for bit"""
    say("case batch with cavity shells")
    X3, T3 = icosphere(3)
    X4, T4 = icosphere(4)
    meshes = [(imperfect(X3, 0.01 * H, seed=31), T3), (imperfect(X4, 0.01 * H, seed=32), T4),
              (imperfect(X3, 0.01 * H, seed=33), T3)]
    base = dict(dt=0.02, dt_max=0.05, adapt_dt_every=1)
    params = [split(**CASES["all of them, shrinking cavity"], **base)[0],
              split(**CASES["box, dynamic friction"], **base)[0],
              dict(pressure_final=-1e-4, **base)]
    b = S.ShellBatch(meshes, params)
    z = np.zeros(sum(X.size for X, _ in meshes))
    b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.apply_initial_conditions()
    b.advance(200)
    for s, ((X, T), p) in enumerate(zip(meshes, params)):
        g = S.ShellSystem(X, T, **p)
        g.set_state(np.zeros(X.size))
        g.apply_initial_conditions()
        g.advance(200)
        ug, vg, _, tg = g.get_state()
        ub, vb, _, tb = b.get_state(s)
        say("  shell", s, "t", tb, tg, "max |du|", np.abs(ub - ug).max())
        assert tb == tg and np.array_equal(ub, ug) and np.array_equal(vb, vg)
        if p.get("cavity_contact"):
            assert np.array_equal(b.cavity_springs(s), g.cavity_springs())
        g.close()
    b.close()


def test_state_file_restart_with_cavity(icosphere, tmp_path):
    """save after 120 steps, load into a fresh context, 120 more steps: the
    same state as 240 steps in one run (the cavity springs travel in the
    state file, ShellSystem.C:2972-2982, 3061-3075)"""
    say("case cavity restart")
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=24)
    case = dict(CASES["all of them, shrinking cavity"], mu_s_cavity=0.5, friction_stop_velocity=0.02)
    kg, _ = split(**case, dt=0.02, dt_max=0.05, adapt_dt_every=1)
    a = S.ShellSystem(X, T, **kg)
    a.set_state(np.zeros(X.size))
    a.apply_initial_conditions()
    a.advance(120)
    path = str(tmp_path / "state.dat")
    a.save_state(path, 120)
    assert np.any(a.cavity_springs() != 0)       # sticking nodes hold springs
    a.advance(120)
    ua, va, _, ta = a.get_state()
    b = S.ShellSystem(X, T, **kg)
    b.load_state(path)
    b.advance(120)
    ub, vb, _, tb = b.get_state()
    say("  t", ta, tb, "rel u", relerr(ub, ua), "rel v", relerr(vb, va))
    assert ta == tb and relerr(ub, ua) < 1e-12 and relerr(vb, va) < 1e-12
    a.close()
    b.close()
