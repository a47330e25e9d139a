"""Measurements on the GPU against the CPU oracle: the moving averages of
measure_properties / print_properties (ShellSystem.C:2714-2898; the
accepted steps' dt-weighted samples, accumulated by the step epilogue) and
the per-element and per-node fields of the .vtu output (curvatures over the
element area, nodal mass and kinetic energy), on a deflating level-3 shell
with self-contact and friction."""
import numpy as np
import pytest

import oracle
import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", ["deflation", "closing sheets"])
def test_properties_and_fields_against_oracle(icosphere, case):
    """deflation: the volume-controlled post-buckling run (constant dt);
    closing sheets: a squashed shell whose sheets close into contact with
    friction under the reference's adaptive dt (unequal sample weights)"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=41)
    z = np.zeros(X.size)
    if case == "deflation":
        kw = dict(E2=0.6, bulk_modulus=0.05, volume_final_ratio=0.85, volume_ramp_time=30.0, dt=0.1, dt_max=0.1,
                  damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=1.0,
                  cell_margin=0.1)
        u0, v0, steps = z, z, 150
    else:
        kw = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1, contact=1, mu_s=0.3, mu_d=0.3,
                  cell_size_mult=1.0, rebuild_trigger=1)
        Y = X.copy()
        Y[:, 2] *= 0.03
        u0 = (Y - X).reshape(-1)
        v0 = np.zeros_like(X)
        v0[:, 2] = -0.02 * np.sign(X[:, 2])
        v0 = v0.reshape(-1)
        steps = 70
    g = S.ShellSystem(X, T, **kw)
    o = oracle.Oracle(X, T, **kw)
    g.set_state(u0, v0)
    o.set_state(u0, v0)
    g.apply_initial_conditions()
    o.init_dynamics()
    g.measure()
    o.measure()
    for chunk in range(3):
        g.advance(steps)
        o.step(steps)
        pg = g.properties()
        po = o.properties()
        say("case properties chunk", chunk, {k: f"{pg[k]:.6g}" for k in ("area", "volume", "Nc_self", "E_kin",
                                                                        "E_self", "kGauss", "kMean")})
        for k, name in enumerate(S.PROPERTY_COLUMNS):
            a, b = pg[name], po[k]
            assert abs(a - b) <= 1e-9 * max(abs(b), 1e-12), (name, a, b)
    if case == "closing sheets":
        assert pg["Nc_self"] > 0 and pg["E_self"] > 0    # the contact columns are exercised
    # fields at the current state, against the oracle evaluated at the same state
    nf, ef = g.fields()
    u, v, _, t = g.get_state()
    o.forces(u, v, t)
    m = o.elem_measure()
    Minv, _ = o.mass()
    for col, ref in ((2, m[:, 7] / m[:, 9]), (3, m[:, 8] / m[:, 9])):
        err = np.max(np.abs(ef[:, col] - ref)) / np.max(np.abs(ref))
        say("  element field", S.ELEM_FIELDS[col], "max rel err", err)
        assert err < 1e-10
    assert np.allclose(nf[:, 4], 1 / Minv[::3], rtol=1e-13, atol=0)
    vv = v.reshape(-1, 3)
    ek = 0.5 * (vv ** 2).sum(1) * nf[:, 4]
    area = m[:, 9]
    node_area = np.zeros(len(X))
    cnt = np.zeros(len(X))
    for e, tri in enumerate(T):
        node_area[tri] += area[e]
        cnt[tri] += 1
    assert np.allclose(nf[:, 2], ek / (node_area / cnt), rtol=1e-9, atol=1e-300)
    assert np.all(nf[:, 0] == 0) and np.all(nf[:, 3] >= 0)
    g.close()
    o.close()


def test_fields_leave_the_trajectory_unchanged(icosphere):
    """a fields evaluation in the middle of a run with sticking and sliding
    contacts changes nothing the run uses (pair distances, friction springs,
    element areas): the run with it equals the run without it bit for bit"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=42)
    kw = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, adapt_dt_every=1, contact=1, mu_s=0.3, mu_d=0.3,
              cell_size_mult=1.0, rebuild_trigger=1)
    Y = X.copy()
    Y[:, 2] *= 0.03
    u0 = (Y - X).reshape(-1)
    v0 = np.zeros_like(X)
    v0[:, 2] = -0.02 * np.sign(X[:, 2])
    out = []
    for with_fields in (False, True):
        g = S.ShellSystem(X, T, **kw)
        g.set_state(u0, v0.reshape(-1))
        g.apply_initial_conditions()
        steps = 0
        while True:                                  # to a step with penetrating contacts
            r = g.advance(10)
            steps += 10
            if r["n_contacts"] > 0 or steps >= 400:
                break
        assert r["n_contacts"] > 0
        if with_fields:
            nf, _ = g.fields()
            assert np.max(nf[:, 3]) > 0              # the contacts' energy shares
        g.advance(80)
        out.append(g.get_state()[0])
        g.close()
    say("case fields mid-run: contacts", r["n_contacts"])
    assert np.array_equal(out[0], out[1])
