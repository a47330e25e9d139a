"""Open meshes on the GPU against the CPU FP64 oracle: the ghost layer's
elements complete the real elements' patches and carry no force, mass or
contact of their own; the boundary conditions of apply_boundary_conditions
(ShellSystem.C:2138-2292), strong and weak, their bc_coords, bc_every and
bc_criterion; the ghost nodes left out of check_increment and the kinetic
energy (ShellSystem.C:2609, 2798).  Bar: states (ghost nodes included)
and reports within 1e-10 relative; a batch of open and closed shells equal
to its shells run alone bit for bit."""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import plate, tube

pytestmark = pytest.mark.gpu

ISO = dict(E1=1.0, E2=1.0, nu=0.3, G12=1.0 / 2.6, rho=1.0)
NAMES = ["clamped", "clamped weak", "pinned", "pinned weak", "free", "free weak"]


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def say(*a):
    print(*a, flush=True)


def open_mesh(kind):
    X, T = plate(8) if kind == "plate" else tube(16, 6)
    Xg, Tg, ng = S.mesh_add_ghosts(X, T)
    return Xg, Tg, ng, len(X)


def stable_dt(Xg, Tg, kw, cfl=0.6):
    """a step within the explicit limit, the weak springs on the light ghost nodes included"""
    o = oracle.Oracle(Xg, Tg, **dict(kw, dt=1.0))
    om = o.cfl_omega_max()
    if kw.get("shell_boundary") in (1, 3, 5):
        om = max(om, np.sqrt(kw.get("bc_stiff_mult", 1.0) * kw["thickness"] / 0.91 * o.mass()[0].max()))
    o.close()
    return cfl / om


def initial_state(Xg, seed):
    """a smooth displacement and velocity field"""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(3, 3))
    u = 0.002 * np.sin(Xg @ c)
    v = 0.001 * np.cos(Xg @ c[::-1])
    return u.reshape(-1), v.reshape(-1)


def run_pair(Xg, Tg, kw, steps, seed=1, keep=None, crit=""):
    g = S.ShellSystem(Xg, Tg, **dict(kw, bc_criterion_fn=crit))
    o = oracle.Oracle(Xg, Tg, **kw)
    if keep is not None:
        be = o.boundary_edges()
        o.bc_keep(keep(0.5 * (Xg[be[:, 1]] + Xg[be[:, 2]])))
    u, v = initial_state(Xg, seed)
    g.set_state(u, v)
    o.set_state(u, v)
    g.apply_initial_conditions()
    o.init_dynamics()
    for k in range(len(steps)):
        rg = g.advance(steps[k])
        ro = o.step(steps[k])
        ug, vg, ag, tg = g.get_state()
        uo, vo, ao, to = o.get_state()
        say("  t", tg, to, "rel u", relerr(ug, uo), "v", relerr(vg, vo), "a", relerr(ag, ao), "ekin", rg["e_kin"],
            ro["e_kin"], "du_max", rg["du_max"], ro["du_max"], "rejected", rg["n_rejected"], ro["n_rejected"])
        assert rg["n_rejected"] == ro["n_rejected"] and abs(tg - to) <= 1e-12 * max(1.0, abs(to))
        assert relerr(ug, uo) < 1e-10 and relerr(vg, vo) < 1e-10 and relerr(ag, ao) < 1e-10
        assert rg["e_kin"] == pytest.approx(ro["e_kin"], rel=1e-10)
        assert rg["du_max"] == pytest.approx(ro["du_max"], rel=1e-10)
    g.close()
    o.close()
    return ug, uo


@pytest.mark.parametrize("bc", range(6))
def test_plate_conditions_equal_the_oracle(bc):
    """a plate of 8 x 8 squares under its own weight from a smooth initial
    state: each condition, the corners' pinned edges whose n1 another edge
    holds at zero among them"""
    say("case plate", NAMES[bc])
    Xg, Tg, ng, nv = open_mesh("plate")
    kw = dict(ISO, thickness=0.05, gravity=3e-3, gravity_dir=1, damping_alpha=0.1, ghost_elems=ng, shell_boundary=bc,
              bc_every=1)
    kw["dt"] = stable_dt(Xg, Tg, kw)
    ug, _ = run_pair(Xg, Tg, kw, [20, 20])
    if bc in (0, 2):
        o = oracle.Oracle(Xg, Tg, **kw)
        be = o.boundary_edges()
        o.close()
        held = be[:, :3] if bc == 0 else be[:, 1:3]
        assert np.abs(ug.reshape(-1, 3)[held.ravel()]).max() == 0


@pytest.mark.parametrize("bc", [2, 4, 5])
def test_tube_under_pressure_with_adaptive_steps(bc):
    """an open tube under a pressure ramp, adaptive dt: rejected attempts
    and states equal"""
    say("case tube", NAMES[bc])
    Xg, Tg, ng, nv = open_mesh("tube")
    kw = dict(ISO, thickness=0.02, pressure_final=1e-4, pressure_ramp_time=2.0, damping_alpha=0.2, ghost_elems=ng,
              shell_boundary=bc, bc_every=1, adapt_dt_every=1, rel_loc_err_min=5e-5, rel_loc_err_max=5e-4)
    kw["dt"] = 4 * stable_dt(Xg, Tg, kw)
    run_pair(Xg, Tg, kw, [15, 25])


def test_coordinates_every_and_criterion():
    """bc_coords 5 (x and z), bc_every 2 and bc_criterion x < 0.5 (the
    product evaluates the expression, the oracle keeps the edges by mask)"""
    say("case bc_coords / bc_every / bc_criterion")
    Xg, Tg, ng, nv = open_mesh("plate")
    kw = dict(ISO, thickness=0.05, gravity=3e-3, gravity_dir=1, damping_alpha=0.1, ghost_elems=ng, shell_boundary=2,
              bc_every=2, bc_coords=5)
    kw["dt"] = stable_dt(Xg, Tg, kw)
    run_pair(Xg, Tg, kw, [15, 16], seed=3)
    kw.update(bc_every=1, bc_coords=7, shell_boundary=0)
    run_pair(Xg, Tg, kw, [20], seed=4, keep=lambda m: m[:, 0] < 0.5, crit="x<0.5")


def test_batch_of_open_and_closed_shells_equals_standalone():
    """a pinned plate, a closed icosphere under pressure and a free tube in
    one batch against each run alone: bit for bit"""
    say("case open and closed batch")
    Xp, Tp, np_, _ = open_mesh("plate")
    Xt, Tt, nt, _ = open_mesh("tube")
    Xs, Ts = S.mesh_icosphere(2)
    meshes = [(Xp, Tp), (Xs, Ts), (Xt, Tt)]
    params = [dict(ISO, thickness=0.05, gravity=3e-3, gravity_dir=1, damping_alpha=0.1, ghost_elems=np_,
                   shell_boundary=2, bc_every=1),
              dict(ISO, thickness=0.02, pressure_final=1e-4, damping_alpha=0.1),
              dict(ISO, thickness=0.02, pressure_final=1e-4, damping_alpha=0.1, ghost_elems=nt, shell_boundary=4,
                   bc_every=1)]
    for (Xk, Tk), p in zip(meshes, params):
        p["dt"] = stable_dt(Xk, Tk, p)
    b = S.ShellBatch(meshes, params)
    states = [initial_state(Xk, 10 + s) for s, (Xk, _) in enumerate(meshes)]
    for s, (u, v) in enumerate(states):
        b.set_state(s, u, v=v, a=np.zeros_like(u), t=0.0)
    b.apply_initial_conditions()
    b.advance(30)
    for s, ((Xk, Tk), p) in enumerate(zip(meshes, params)):
        g = S.ShellSystem(Xk, Tk, **p)
        g.set_state(*states[s])
        g.apply_initial_conditions()
        g.advance(30)
        ug, vg, ag, tg = g.get_state()
        ub, vb, ab, tb = b.get_state(s)
        say("  shell", s, "t", tb, tg, "max |du|", np.abs(ub - ug).max())
        assert tb == tg and np.array_equal(ub, ug) and np.array_equal(vb, vg) and np.array_equal(ab, ag)
        g.close()
    b.close()


@pytest.mark.parametrize("contact", [1, 2])
def test_contact_on_an_open_tube(contact):
    """an open tube with pinned ends squashed flat until its walls touch:
    self-contact with friction on an open mesh, linear and smooth (the ghost
    elements take no part in the cell lists or close sets, ShellSystem.C:
    664-674, 700, 1974, 2035; the smooth contact's limit surfaces of the
    boundary elements run over ghost nodes).  Linear: pair lists bit for
    bit and states within 1e-10 over two chunks of steps.  Smooth: the
    flattened walls are straight along the tube's axis, so two facing
    elements' closest points are not unique (only their distance is) and
    each side's search stops at another point of the minimising set; the
    contact forces then differ at 1e-6 and the adaptive step carries that
    into dt.  So the pair lists and distances at the start are compared (1e-8),
    and after one chunk the pair lists and contact counts exactly, u
    within 1e-5 and t within 1e-4 (measured 1e-6 and 5e-5)."""
    say("case contact on an open tube", contact)
    X, T = tube(24, 6, r=0.5, length=1.0)
    Xg, Tg, ng = S.mesh_add_ghosts(X, T)
    kw = dict(ISO, thickness=0.02, damping_alpha=0.1, ghost_elems=ng, shell_boundary=2, bc_every=1, contact=contact,
              mu_s=0.3, mu_d=0.3, cell_size_mult=1.0)
    kw["dt"] = stable_dt(Xg, Tg, kw, cfl=0.4)
    # adaptive steps: the flattened folds start far from equilibrium
    kw.update(adapt_dt_every=1, dt_max=kw["dt"], rel_loc_err_min=5e-5, rel_loc_err_max=5e-4)
    u = np.zeros_like(Xg)
    u[:, 0] = -Xg[:, 0] * 0.97                   # the walls 0.03 apart
    v = np.zeros_like(Xg)
    v[:, 0] = -0.02 * np.sign(Xg[:, 0])
    g = S.ShellSystem(Xg, Tg, **kw)
    o = oracle.Oracle(Xg, Tg, **kw)
    g.set_state(u.reshape(-1), v.reshape(-1))
    o.set_state(u.reshape(-1), v.reshape(-1))
    g.apply_initial_conditions()
    o.init_dynamics()
    if contact == 2:
        assert np.array_equal(np.concatenate(g.pairs()), np.concatenate(o.pairs()))
        d2g, d2o = g.pair_state()["d2"], o.pair_state()[0]
        say("  start: pairs", len(d2o), "max rel d2 diff", np.max(np.abs(d2g - d2o) / np.maximum(d2o, 1e-300)))
        assert np.allclose(d2g, d2o, rtol=1e-8, atol=0)
    for k in range(2 if contact == 1 else 1):
        rg = g.advance(20)
        ro = o.step(20)
        ug, vg, _, tg = g.get_state()
        uo, vo, _, to = o.get_state()
        e1g, e2g = g.pairs()
        e1o, e2o = o.pairs()
        say("  t", tg, to, "pairs", len(e1g), len(e1o), "contacts", rg["n_contacts"], ro["n_contacts"], "rel u",
            relerr(ug, uo), "v", relerr(vg, vo))
        assert np.array_equal(e1g, e1o) and np.array_equal(e2g, e2o) and len(e1g) > 0
        assert rg["n_contacts"] == ro["n_contacts"] and rg["n_rebuilds"] == ro["n_rebuilds"]
        assert rg["n_rejected"] == ro["n_rejected"]
        if contact == 1:
            assert abs(tg - to) <= 1e-12 * max(1.0, abs(to))
            assert relerr(ug, uo) < 1e-10 and relerr(vg, vo) < 1e-10
        else:
            assert abs(tg - to) <= 1e-4 * abs(to) and relerr(ug, uo) < 1e-5
    assert ro["n_contacts"] > 0
    g.close()
    o.close()
