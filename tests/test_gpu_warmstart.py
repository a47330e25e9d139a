"""Warm-start mechanics of the batch: all-shell state transfer, the
re-initialisation of a subset of shells (the other shells' runs go on
untouched; the re-initialised shell restarts as if fresh), a caller-given
reference volume V0, and the relaxation judging a warm state.
"""
import numpy as np
import pytest

import shell_amd as S
from conftest import imperfect
from shell_amd import warmstart as W

pytestmark = pytest.mark.gpu

H = 0.01
PCR = 2 * 1.0 * H * H / np.sqrt(3 * (1 - 0.3 ** 2))


def say(*a):
    print(*a, flush=True)


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _shells(icosphere):
    X3, T3 = icosphere(3)
    meshes = [(imperfect(X3, 0.01 * H, seed=1), T3), (imperfect(X3, 0.01 * H, seed=2), T3),
              (imperfect(X3, 0.02 * H, seed=3), T3)]
    params = [
        dict(pressure_final=-1.5 * PCR, pressure_ramp_time=20.0, dt=0.1, damping_alpha=0.05),
        dict(E2=0.3, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=30.0, dt=0.1, dt_max=0.1,
             damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.2, mu_d=0.2, cell_size_mult=4.0,
             cell_margin=0.1),
        dict(E2=0.8, thickness=0.015, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=20.0, dt=0.1,
             dt_max=0.1, damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.4, mu_d=0.4,
             cell_size_mult=4.0, cell_margin=0.1),
    ]
    return meshes, params


def _fresh(b):
    z = np.zeros(3 * int(b.nv.sum()))
    b.set_state(S.ALL_SHELLS, z, v=z, a=z, t=0.0)


def test_all_shell_state_roundtrip(icosphere):
    meshes, params = _shells(icosphere)
    say("case set/get state of all shells at once,", len(meshes), "shells")
    b = S.ShellBatch(meshes, params)
    n = 3 * int(b.nv.sum())
    rng = np.random.default_rng(5)
    u, v, a = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    b.set_state(S.ALL_SHELLS, u, v, a, t=1.5)
    off = 0
    for s in range(b.S):
        m = 3 * int(b.nv[s])
        us, vs, as_, ts = b.get_state(s)
        assert np.array_equal(us, u[off:off + m]) and np.array_equal(vs, v[off:off + m])
        assert np.array_equal(as_, a[off:off + m]) and ts == 1.5
        off += m
    ua, va, aa, ta = b.get_state(S.ALL_SHELLS)
    assert np.array_equal(ua, u) and np.array_equal(va, v) and np.array_equal(aa, a) and ta == 1.5


def test_init_subset_leaves_other_shells(icosphere):
    """re-initialising shell 1 mid-run: shell 0 (no contact: no atomics) goes
    on bit-exactly, shell 2 within the FP64 bar; shell 1 restarts exactly as
    a fresh start"""
    meshes, params = _shells(icosphere)
    say("case init_shells([1]) after 150 steps vs an uninterrupted run")
    a = S.ShellBatch(meshes, params)
    ref = S.ShellBatch(meshes, params)
    for b in (a, ref):
        _fresh(b)
        b.apply_initial_conditions()
    a.advance(150)
    ref.advance(300)
    z = np.zeros(3 * int(a.nv[1]))
    a.set_state(1, z, v=z, a=z, t=0.0)
    a.init_shells([1])
    ra = a.advance(150)
    u0a, v0a, _, t0a = a.get_state(0)
    u0r, v0r, _, t0r = ref.get_state(0)
    assert np.array_equal(u0a, u0r) and np.array_equal(v0a, v0r) and t0a == t0r
    u2a, _, _, _ = a.get_state(2)
    u2r, _, _, _ = ref.get_state(2)
    e2 = relerr(u2a, u2r)
    say(f"  shell 2 rel. diff {e2:.2e}")
    assert e2 < 1e-10
    # shell 1 after its restart = a fresh 150-step run
    fresh = S.ShellBatch(meshes, params)
    _fresh(fresh)
    fresh.apply_initial_conditions()
    rf = fresh.advance(150)
    u1a, _, _, t1a = a.get_state(1)
    u1f, _, _, t1f = fresh.get_state(1)
    e1 = relerr(u1a, u1f)
    say(f"  shell 1 restart vs fresh rel. diff {e1:.2e}, steps {ra[1]['n_steps']} vs {rf[1]['n_steps']}")
    assert ra[1]["n_steps"] == rf[1]["n_steps"] == 150 and t1a == t1f
    assert e1 < 1e-10


def test_given_volume0_equals_own(icosphere):
    """V0 given as the state's own volume reproduces the default init bit-exactly"""
    meshes, params = _shells(icosphere)
    meshes, params = meshes[:1], [dict(params[2], contact=0)]
    say("case init_shells(volume0 = own volume) vs default init")
    out = []
    for given in (False, True):
        b = S.ShellBatch(meshes, params)
        _fresh(b)
        b.apply_initial_conditions()
        if given:
            V = b.reports()[0]["volume"]
            b.init_shells([0], volume0=[V])
        b.advance(200)
        out.append(b.get_state(0)[0])
    assert np.array_equal(out[0], out[1])


def test_warm_start_from_equilibrium(icosphere):
    """an equilibrium fed back as a warm start at full load, with V0 the
    undeformed volume, is accepted within a few steps; with V0 taken from the
    deformed state instead (the plain init's rule) it is not an equilibrium"""
    X, T = icosphere(3)
    X = imperfect(X, 0.01 * H, seed=7)
    prm = dict(E2=0.5, bulk_modulus=0.05, volume_final_ratio=0.9, volume_ramp_time=60.0, dt=0.1, dt_max=0.1,
               damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.3, mu_d=0.3, cell_size_mult=4.0,
               cell_margin=0.1)
    say("case L3 equilibrium reused as warm start")
    b = S.ShellBatch([(X, T)], [prm])
    _fresh(b)
    b.apply_initial_conditions()
    V_ref = b.reports()[0]["volume"]
    r = None
    for _ in range(20):
        r = b.advance(2000, tol=1e-7, stop_on_converged=True)[0]
        if r["converged_step"] >= 0:
            break
    assert r["converged_step"] >= 0, r
    plain_steps = r["converged_step"]
    u_eq = b.get_state(0)[0]
    u_al = W.align([X], u_eq.reshape(1, -1))[0]
    z = np.zeros_like(u_eq)
    b.set_state(0, u_al, v=z, a=z, t=60.0)
    b.init_shells([0], volume0=[V_ref])
    r0 = b.reports()[0]
    say(f"  plain steps {plain_steps}, warm initial residual {r0['residual_deform']:.2e}")
    assert r0["problem"] == 0
    assert r0["residual_deform"] < 1e-5
    r = b.advance(2000, tol=1e-6, stop_on_converged=True)[0]
    say(f"  warm converged at step {r['converged_step']}")
    assert 0 <= r["converged_step"] <= 50
    b.set_state(0, u_al, v=z, a=z, t=60.0)
    b.init_shells([0])
    assert b.reports()[0]["residual_deform"] > 1e-3


def test_trained_surrogate_gains_over_plain_start():
    """the network trained (on the GPU, by the library's kernels) on the
    committed equilibria of this solver seeds the relaxation of unseen shells:
    every shell reaches the 1e-6 residual, in fewer total steps than from the
    plain start (fallbacks counted with all their steps)"""
    import os
    from shell_amd.surrogate import Surrogate
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    d = np.load(os.path.join(root, "shell_amd", "data", "warm_L4_4096.npz"))
    conv = d["converged"]
    sg = Surrogate(d["umean"], d["basis"][:64].astype(np.float64), seed=1)
    for ep, lr in ((100, 1e-3), (50, 3e-4), (50, 1e-4)):
        loss = sg.fit(d["params"][conv], d["coeffs"][:, :64].astype(np.float64), epochs=ep, batch=256, lr=lr)
    say(f"case surrogate trained on {int(conv.sum())} equilibria, loss {loss:.3f}; 32 unseen level-4 shells")
    assert loss < 0.6
    P = W.draw_params(32, 12)
    b, meshes = W.build(P)
    vref = W.reference_volumes(b)
    rp, wall_p = W.plain_solve(b, 1e-6, 60000, 5000)
    rw, wall_w, info = W.warm_solve(b, sg, P, vref, 1e-6, 60000, 5000)
    sp = [r["converged_step"] for r in rp]
    sw = [r["n_steps"] + (W.QUENCH_STEPS + W.PROBE_STEPS if i in info["fallback"] else 0) for i, r in enumerate(rw)]
    say(f"  plain steps {sum(sp)} ({wall_p:.1f} s), warm {sum(sw)} ({wall_w:.1f} s), fallbacks {len(info['fallback'])}")
    assert all(s >= 0 for s in sp) and all(r["converged_step"] >= 0 for r in rw)
    assert sum(sw) < sum(sp)


def test_capacity_error_stops_only_its_shell(icosphere):
    """a shell whose contact rows overflow (a squashed level-3 shell with a
    candidate radius that lists up to 80 pairs per row; rows hold 64) stops
    with capacity_err set while the other shell of the batch steps on; the
    call reports SHX_E_CAPACITY; re-initialising the shell from a state whose
    rows fit clears the error"""
    X, T = icosphere(3)
    X0 = imperfect(X, 0.01 * H, seed=3)
    pa = dict(E2=0.5, bulk_modulus=0.05, volume_final_ratio=0.95, volume_ramp_time=30.0, dt=0.1, dt_max=0.1,
              damping_alpha=0.05, volume_at_prediction=1, contact=1, mu_s=0.2, mu_d=0.2, cell_size_mult=4.0,
              cell_margin=0.1)
    pb = dict(dt=0.02, dt_max=0.02, damping_alpha=0.05, contact=1, cell_size_mult=1.0, cell_margin=1.0)
    say("case row overflow in shell 1 of 2")
    b = S.ShellBatch([(X0, T), (X, T)], [pa, pb])
    ua = np.zeros(3 * len(X))
    Y = X.copy()
    Y[:, 2] *= 0.01
    ub = (Y - X).reshape(-1)
    z = np.zeros(3 * len(X))
    b.set_state(0, ua, v=z, a=z, t=0.0)
    b.set_state(1, ub, v=z, a=z, t=0.0)
    with pytest.raises(S.ShellError) as ei:
        b.apply_initial_conditions()
    assert ei.value.code == S.E_CAPACITY
    r = b.reports()
    say("  capacity_err", r[0]["capacity_err"], r[1]["capacity_err"])
    assert r[0]["capacity_err"] == 0 and r[1]["capacity_err"] != 0
    with pytest.raises(S.ShellError) as ei:
        b.advance(40)
    assert ei.value.code == S.E_CAPACITY
    r = b.reports()
    assert r[0]["n_steps"] == 40 and r[1]["n_steps"] == 0
    b.set_state(1, z, v=z, a=z, t=0.0)
    b.init_shells([1])
    r = b.advance(40)
    say("  after re-init: steps", r[0]["n_steps"], r[1]["n_steps"], "capacity_err", r[1]["capacity_err"])
    assert r[1]["capacity_err"] == 0 and r[1]["n_steps"] == 40 and r[0]["n_steps"] == 80
