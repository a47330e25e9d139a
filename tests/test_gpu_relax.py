"""Relaxation controls of the batch on the GPU: kinetic damping (velocities
set to zero at kinetic-energy peaks once the load is complete) reaches the
equilibrium the configured damping reaches, in fewer steps, and is inert
while the load still ramps or when off; parked shells sit out the steps
untouched while the others run, and start when initialised; the streaming
solve of the warm-start module runs a queue of shells through fewer slots.
"""
import numpy as np
import pytest

import shell_amd as S
from conftest import imperfect
from shell_amd import warmstart as W

pytestmark = pytest.mark.gpu

H = 0.01


def say(*a):
    print(*a, flush=True)


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _batch(icosphere, n=3, friction=0.0, ratio=0.9):
    """deflated level-3 shells (no friction by default: with friction the
    equilibrium reached depends on the path)"""
    X3, T3 = icosphere(3)
    meshes = [(imperfect(X3, 0.01 * H, seed=10 + k), T3) for k in range(n)]
    params = [dict(E2=0.4 + 0.2 * k, bulk_modulus=0.05, volume_final_ratio=ratio - 0.01 * k * (1 - ratio),
                   volume_ramp_time=30.0, dt=0.1, dt_max=0.1, damping_alpha=0.05, volume_at_prediction=1, contact=1,
                   mu_s=friction, mu_d=friction, cell_size_mult=4.0, cell_margin=0.1, rebuild_trigger=1)
              for k in range(n)]
    b = S.ShellBatch(meshes, params)
    b.X = [X for X, _ in meshes]
    return b


def _rest(b):
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)
    b.apply_initial_conditions()


def _solve(b, tol=1e-6, chunk=2000, max_steps=40000):
    done = 0
    while done < max_steps:
        reps = b.advance(chunk, tol=tol, stop_on_converged=True)
        done += chunk
        if all(r["converged_step"] >= 0 for r in reps):
            break
    return reps


def _deform(b, s):
    """the shell's displacement with its rigid motion removed (the residual
    leaves the rigid-body modes free: runs may drift apart in them)"""
    return W.align([b.X[s]], b.get_state(s)[0][None])[0]


def test_kinetic_damping_reaches_the_same_equilibrium(icosphere):
    """before buckling (V/V0 = 0.995) the equilibrium is unique up to a
    rigid motion: kinetic damping reaches the one the configured damping
    reaches, in fewer steps"""
    say("case kinetic damping vs configured damping, 3 level-3 shells before buckling")
    b = _batch(icosphere, ratio=0.995)
    _rest(b)
    ref = _solve(b)
    U_ref = [_deform(b, s) for s in range(b.S)]
    assert all(r["converged_step"] >= 0 and r["n_kd"] == 0 for r in ref)
    b.set_kinetic_damping(True)
    _rest(b)
    kd = _solve(b)
    for s in range(b.S):
        u = _deform(b, s)
        say(f"  shell {s}: steps {ref[s]['converged_step']} -> {kd[s]['converged_step']}, resets {kd[s]['n_kd']}, "
            f"rel u {relerr(u, U_ref[s]):.2e}")
        assert kd[s]["converged_step"] >= 0 and kd[s]["n_kd"] > 0
        assert relerr(u, U_ref[s]) < 1e-4
    assert sum(r["converged_step"] for r in kd) < sum(r["converged_step"] for r in ref)
    b.close()


def test_kinetic_damping_converges_post_buckled_shells(icosphere):
    """after buckling (V/V0 about 0.9) equilibria are not unique and the two
    relaxations may end in different ones; the kinetically damped end state
    is an equilibrium: re-initialised from it (v = 0, V0 of the undeformed
    shell), its residual is still within the tolerance"""
    say("case kinetic damping, post-buckled shells")
    b = _batch(icosphere)
    V0 = W.reference_volumes(b)
    b.set_kinetic_damping(True)
    _rest(b)
    kd = _solve(b)
    U = b.get_state(S.ALL_SHELLS)[0]
    t = [b.get_state(s)[3] for s in range(b.S)]
    off = 0
    for s in range(b.S):
        m = 3 * int(b.nv[s])
        z = np.zeros(m)
        b.set_state(s, U[off:off + m], v=z, a=z, t=t[s])
        off += m
    b.init_shells(np.arange(b.S), volume0=V0)
    for s, r in enumerate(b.reports()):
        say(f"  shell {s}: converged at {kd[s]['converged_step']}, resets {kd[s]['n_kd']}, "
            f"V/V0 {kd[s]['volume'] / V0[s]:.4f}, residual of the end state {r['residual_deform']:.2e}")
        assert kd[s]["converged_step"] >= 0 and kd[s]["volume"] / V0[s] < 0.95
        assert r["residual_deform"] <= 2e-6
    b.close()


def test_kinetic_damping_waits_for_the_load_and_can_be_turned_off(icosphere):
    say("case kinetic damping inert during the ramp")
    b = _batch(icosphere, n=2, friction=0.2)
    _rest(b)
    b.advance(200)                       # t = 20 < the ramp's 30
    ua = b.get_state(S.ALL_SHELLS)[0]
    b.set_kinetic_damping(True)
    _rest(b)
    reps = b.advance(200)
    assert all(r["n_kd"] == 0 for r in reps)
    assert np.array_equal(b.get_state(S.ALL_SHELLS)[0], ua)
    reps = b.advance(400)                # past the ramp: resets start
    assert all(r["n_kd"] > 0 for r in reps)
    b.set_kinetic_damping([1, 0])        # per shell
    _rest(b)
    reps = b.advance(600)
    assert reps[0]["n_kd"] > 0 and reps[1]["n_kd"] == 0
    with pytest.raises(S.ShellError):
        b.set_kinetic_damping([2, 0])
    b.close()


def test_parked_shells_sit_out(icosphere):
    say("case parked shells")
    b = _batch(icosphere, n=3, friction=0.2)
    _rest(b)
    b.advance(50)
    before = [b.get_state(s) for s in range(b.S)]
    b.park([1])
    reps = b.advance(100)
    assert reps[0]["n_steps"] == 150 and reps[2]["n_steps"] == 150 and reps[1]["n_steps"] == 50
    u1, v1, a1, t1 = b.get_state(1)
    assert np.array_equal(u1, before[1][0]) and np.array_equal(v1, before[1][1]) and t1 == before[1][3]
    # initialising the parked shell starts it again (fresh, from its own state)
    b.init_shells([1])
    reps = b.advance(100)
    assert reps[1]["n_steps"] == 100 and reps[0]["n_steps"] == 250
    b.close()


def test_stream_solve_through_fewer_slots(icosphere):
    """a queue of 4 shells through 2 slots: every shell converges, the
    equilibria equal those of the whole batch solved at once"""
    say("case streaming solve, 4 shells through 2 slots")
    b = _batch(icosphere, n=4)
    _rest(b)
    full = _solve(b)
    U_full = [b.get_state(s)[0] for s in range(b.S)]
    reps, wall, info = W.stream_solve(b, None, None, 2, None, tol=1e-6, max_steps=40000, chunk=1000)
    for s in range(b.S):
        say(f"  shell {s}: converged at {reps[s]['converged_step']} (whole batch {full[s]['converged_step']})")
        assert reps[s]["converged_step"] == full[s]["converged_step"]
        assert np.array_equal(b.get_state(s)[0], U_full[s])
    assert info["fallback"] == []
    b.close()
