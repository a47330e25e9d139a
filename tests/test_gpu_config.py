"""Time functions from a config.cfg on the GPU: the pressure and
volume_growth expressions (muparser dialect, compiled at shx_create and
evaluated once per shell and attempt by the element epilogue) drive the
same arithmetic as the linear ramps they can express, so a run driven by
the expression equals the run driven by the ramp bit for bit (the ramp runs
are the ones checked against the oracle), and a config read by the library
creates and steps the same shell as the equivalent explicit parameters."""
import numpy as np
import pytest

import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))


def say(*a):
    print(*a, flush=True)


def _run(X, T, n, **p):
    g = S.ShellSystem(X, T, **p)
    g.set_state(np.zeros(X.size))
    g.apply_initial_conditions()
    r = g.advance(n)
    u, v, _, t = g.get_state()
    g.close()
    return u, v, r


@pytest.mark.parametrize("kind", ["pressure", "volume"])
def test_expression_equals_ramp(icosphere, kind):
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=31)
    base = dict(dt=0.1, damping_alpha=0.05)
    if kind == "pressure":
        ramp = dict(base, pressure_final=-2 * PCR, pressure_ramp_time=5.0)
        expr = dict(base, pressure_fn=f"{float(-2 * PCR)!r}*min(t/5,1)", load_end_time=5.0)
    else:
        ramp = dict(base, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=8.0, contact=1,
                    cell_size_mult=1.0)
        expr = dict(base, bulk_modulus=0.05, volume_growth_fn="1-(1-0.9)*min(t/8,1)", load_end_time=8.0,
                    contact=1, cell_size_mult=1.0)
    say("case time function", kind)
    ua, va, ra = _run(X, T, 120, **ramp)
    ub, vb, rb = _run(X, T, 120, **expr)
    say("  |u| ramp", np.linalg.norm(ua), "expression", np.linalg.norm(ub), "volume", ra["volume"], rb["volume"])
    assert np.linalg.norm(ua) > 0
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
    assert ra["volume"] == rb["volume"] and ra["e_tens"] == rb["e_tens"]


def test_config_creates_the_same_shell(icosphere):
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=32)
    text = f"""shell_mass_density = 1
shell_young_modulus = 1
shell_young_modulus2 = 0.5
shell_poisson_ratio = 0.3
shell_shear_modulus = 0.35
thickness = {H}
damping_alpha = 0.05
pressure = '{float(-1.5 * PCR)!r}*(t<4 ? t/4 : 1)'   # a quoted GetPot value may hold spaces
shell_shell_contact = none
shell_cavity_contact = false
newmark_beta = 0
dt = 0.1
adapt_dt_every = 0
"""
    p, _, _ = S.read_config(text)
    say("case config-created shell")
    ua, va, ra = _run(X, T, 80, **p)
    explicit = dict(E1=1.0, E2=0.5, nu=0.3, G12=0.35, thickness=H, rho=1.0, damping_alpha=0.05,
                    pressure_final=-1.5 * PCR, pressure_ramp_time=4.0, newmark_beta=0.0, dt=0.1)
    ub, vb, rb = _run(X, T, 80, **explicit)
    rel = np.linalg.norm(ua - ub) / np.linalg.norm(ub)
    say("  rel diff", rel)
    # t/4 and min(t/4, 1) agree for t <= 4, and both are exactly 1 beyond
    assert np.array_equal(ua, ub)
