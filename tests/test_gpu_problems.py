"""Run-time problems and when a shell stops (check_for_problems,
ShellSystem.C:2901-2952; severities ShellSystem.h:48-49): a problem of
severity STOP_SIM or above ends the shell's run in the step where it occurs,
and that shell never counts as converged; with stop_on_problem any problem
ends it.  Every case here is forced from the host through the state it sets
(no kernel is made to fault): touching triangles, a node given a velocity
that carries it far out, and a shell warm-started at an FP64 equilibrium in
the FP32 split mode."""
import numpy as np
import pytest

import shell_amd as S
from conftest import imperfect

pytestmark = pytest.mark.gpu

H = 0.01


def say(*a):
    print(*a, flush=True)


def _contact_shell(icosphere, **kw):
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=11)
    p = dict(bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=0.0, dt=0.05, dt_max=0.05,
             damping_alpha=0.05, contact=1, cell_size_mult=1.0, rebuild_trigger=1)
    p.update(kw)
    return X, T, p


def test_touching_triangles_stop_with_contact_problem(icosphere):
    """two vertices on opposite sides of the shell moved onto one point: the
    triangles around them touch (distance 0, ShellShellContact.C:45), the
    initial contact search registers CONTACT_PROBLEM (and the log-barrier
    force at distance 0 is not finite: NOT_FINITE), the shell takes no step
    and does not converge, even with a tolerance every state meets"""
    X, T, p = _contact_shell(icosphere)
    a = int(np.argmax(X[:, 2]))
    b = int(np.argmin(X[:, 2]))
    target = np.array([0.0, 0.0, -1.25])
    u = np.zeros_like(X)
    u[a] = target - X[a]
    u[b] = target - X[b]
    # the positions X + u the kernels form must meet exactly
    assert np.array_equal(X[a] + u[a], X[b] + u[b])
    say("case touching triangles: vertices", a, b)
    g = S.ShellSystem(X, T, **p)
    g.set_state(u.reshape(-1))
    g.apply_initial_conditions()
    r = g.advance(50, tol=1e30, stop_on_converged=True)
    say("  problem", S.PROBLEMS[r["problem"]], "steps", r["n_steps"], "converged", r["converged_step"])
    assert r["problem"] >= 6
    assert r["n_steps"] == 0 and r["converged_step"] == -1
    assert S.stopped(r)
    g.close()


def test_far_node_stops_with_too_many_cells(icosphere):
    """a node given a velocity that carries it far out in one predictor: its
    box makes the cell grid exceed n_cells_max (ShellSystem.C:1957-1960),
    the rebuild aborts (the old list is kept and the rebuild not counted)
    and the shell stops in that attempt: no further step is accepted (step
    counts run from the shell's init)"""
    X, T, p = _contact_shell(icosphere)
    g = S.ShellSystem(X, T, **p)
    z = np.zeros(X.size)
    g.set_state(z)
    g.apply_initial_conditions()
    r0 = g.advance(20)
    assert r0["problem"] < S.STOP_SIM and r0["n_steps"] == 20
    u, v, a, t = g.get_state()
    v = v.copy()
    v[3 * 7:3 * 7 + 3] = 1e3  # 50 length units along each axis in one step of dt = 0.05
    g.set_state(u, v, a, t)
    say("case far node: rebuilds before", r0["n_rebuilds"])
    r = g.advance(20, tol=1e30, stop_on_converged=True)
    say("  problem", S.PROBLEMS[r["problem"]], "steps", r["n_steps"], "rebuilds", r["n_rebuilds"])
    assert r["problem"] == 4
    assert r["n_steps"] == r0["n_steps"] and r["converged_step"] == -1
    assert r["n_rebuilds"] == r0["n_rebuilds"]
    # the shell stays stopped
    r2 = g.advance(5)
    assert r2["n_steps"] == r0["n_steps"] and r2["problem"] == 4
    g.close()


@pytest.mark.parametrize("stop_on_problem", [0, 1])
def test_large_increment_warning_and_stop_on_problem(icosphere, stop_on_problem):
    """an increment above h/2 with contact is LARGE_DISP, a warning below
    STOP_SIM (ShellSystem.C:2639-2640): the run goes on, unless
    stop_on_problem is set, which ends it after that step
    (ShellSystem.C:2952)"""
    X, T, p = _contact_shell(icosphere, stop_on_problem=stop_on_problem)
    g = S.ShellSystem(X, T, **p)
    z = np.zeros(X.size)
    g.set_state(z)
    g.apply_initial_conditions()
    u, v, a, t = g.get_state()
    v = v.copy()
    v[3 * 7:3 * 7 + 3] = 0.2   # 0.01 = h per step of dt = 0.05: above h/2
    g.set_state(u, v, a, t)
    r = g.advance(10)
    say("case LARGE_DISP stop_on_problem", stop_on_problem, ": problem", S.PROBLEMS[r["problem"]], "steps",
        r["n_steps"])
    assert r["problem"] == 2
    assert r["n_steps"] == (1 if stop_on_problem else 10)
    g.close()


def test_fp32_warm_start_converges_in_fp64(icosphere):
    """a shell started at its FP64 equilibrium (with the undeformed V0) in
    the FP32 split mode: FP32 forces leave a residual floor above the
    tolerance, so the solve takes FP32 steps first and converges only after
    the switch to FP64"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=2024)
    kw = dict(bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=60.0, dt=0.1175, dt_max=0.1175,
              damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.3, mu_d=0.3, cell_margin=0.1)
    b = S.ShellBatch([(X, T)], [kw])
    z = np.zeros(X.size)
    b.set_state(0, z)
    b.apply_initial_conditions()
    V0 = b.reports()[0]["volume"]
    r = b.advance(30000, tol=1e-6, stop_on_converged=True)[0]
    assert r["converged_step"] > 0
    u, _, _, t = b.get_state(0)
    say("case FP32 warm start at the FP64 equilibrium (t =", t, ", V0 =", V0, ")")
    b.set_state(0, u, z, z, t)
    b.init_shells([0], volume0=[V0], precision=S.FP32_SPLIT)
    r = b.advance(20000, tol=1e-6, stop_on_converged=True, precision=S.FP32_SPLIT)[0]
    say("  FP32 steps", r["n_steps_fp32"], "converged at", r["converged_step"], "residual", r["residual_deform"])
    assert r["n_steps_fp32"] > 0
    assert r["converged_step"] > r["n_steps_fp32"]
    b.close()
